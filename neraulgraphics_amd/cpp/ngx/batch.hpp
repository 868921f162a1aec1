// ngx/batch.hpp -- a C++ BatchRenderer over the C ABI of include/ngx_backend.h, header only.
//
// It mirrors the reference's C++ class of the same name (include/neraulgraphics/batch.h:52-221) for the
// batched-world path: the same methods, argument meaning and conventions -- matrices, body poses and cameras
// from host or device memory, results in the renderer's own buffers (BatchFrame, RayHits) or copied out
// (render_host, cast_rays_host), every call finished on return -- and the same error classes:
// std::invalid_argument for bad input, checked before anything changes (src/batch.cpp:91-192), and
// ngx::DeviceError, the counterpart of the reference's CudaError (include/neraulgraphics/renderer.h:32-39),
// for device failures. The RGB picture is the reference's classic one; textures and environment maps are not
// supported.
//
// Link with -lngx (neraulgraphics_amd/libngx.so) and include include/ and this directory.
#pragma once
#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ngx_backend.h"

namespace ngx {

enum class Memory { Host, Device };  // include/neraulgraphics/renderer.h:49

enum Output : unsigned {  // include/neraulgraphics/renderer.h:53-58
    OUTPUT_RGB = NGX_OUTPUT_RGB,
    OUTPUT_DEPTH = NGX_OUTPUT_DEPTH,
    OUTPUT_NORMAL = NGX_OUTPUT_NORMAL,
    OUTPUT_OBJECT_ID = NGX_OUTPUT_OBJECT_ID,
};

constexpr int OBJECT_NONE = NGX_OBJECT_NONE, OBJECT_SPHERE = NGX_OBJECT_SPHERE, OBJECT_FLOOR = NGX_OBJECT_FLOOR,
              OBJECT_INSTANCE = NGX_OBJECT_INSTANCE;

// A device failure, with the backend's status.
class DeviceError : public std::runtime_error {
public:
    DeviceError(const ngx_status status, const std::string &what) : std::runtime_error(what), status_(status) {}
    ngx_status status() const { return status_; }

private:
    ngx_status status_;
};

// Something the reference does that this backend does not (textured meshes, environment maps, the path tracer).
class Unsupported : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

inline void check(const ngx_status s) {
    if (s == NGX_OK) return;
    const std::string what = ngx_last_error();
    if (s == NGX_INVALID_ARGUMENT) throw std::invalid_argument(what);
    if (s == NGX_UNSUPPORTED) throw Unsupported(what);
    throw DeviceError(s, what);
}

struct Vec3 {
    float x, y, z;
};

struct Material {  // the classic fields of include/neraulgraphics/scene.h:35-40 (python/bindings.cpp:393-395 defaults)
    Vec3 diffuse_color = {0.8f, 0.8f, 0.8f};
    float diffuse_albedo = 0.9f, specular_albedo = 0.1f, reflectivity = 0.0f, specular_exponent = 50.0f;
};

struct Sphere {  // include/neraulgraphics/scene.h:63-67
    Vec3 center;
    float radius;
    Material material{};
};

struct Plane {  // include/neraulgraphics/scene.h:96-103; half_extent_x 0 means no floor
    float y = 0.0f, half_extent_x = 0.0f, z_min = 0.0f, z_max = 0.0f, checker_scale = 0.5f;
    Material material{{1.0f, 1.0f, 1.0f}, 1.0f, 0.0f, 0.0f, 1.0f};
};

struct Light {  // include/neraulgraphics/scene.h:78-92
    Vec3 position;
    float intensity = 1.0f, radius = 0.0f, falloff = 0.0f;
    Vec3 direction = {0.0f, 0.0f, -1.0f};
    float spot_angle = 0.0f, spot_softness = 0.25f;
    Vec3 color = {1.0f, 1.0f, 1.0f};
};

struct Environment {  // include/neraulgraphics/scene.h:109-121, the gradient only
    Vec3 up = {0.0f, 1.0f, 0.0f};
    Vec3 color_up = {0.10f, 0.26f, 0.60f};
    Vec3 color_down = {0.35f, 0.52f, 0.70f};
    float intensity = 1.0f;
};

struct Mesh {  // include/neraulgraphics/scene.h: vertices and triangles, optional per-vertex normals
    std::vector<float> vertices;    // 3 per vertex
    std::vector<uint32_t> indices;  // 3 per triangle
    std::vector<float> normals;     // 3 per vertex, or empty
};

struct Instance {  // a mesh placed by the top three rows of a 4x4 transform
    int mesh = 0;
    float transform[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    Material material{};
};

struct BodyBinding {  // include/neraulgraphics/batch.h:52-57
    int bodies_per_world = 0;
    std::vector<int> body_of_instance;  // per instance: 0..bodies_per_world-1, or -1 for fixed in the world
    std::vector<float> local_pose;      // per instance: x, y, z, then a quaternion x, y, z, w
    std::vector<float> scale;           // per instance: x, y, z
};

struct BatchFrame {  // include/neraulgraphics/batch.h:126-136; always device memory here
    Memory memory;
    int device_id, views, width, height;
    const uint8_t *rgb;
    const float *depth;
    const float *normal;
    const int32_t *object_id;
};

struct HostBatchFrame {  // include/neraulgraphics/batch.h:138-146
    int views = 0, width = 0, height = 0;
    std::vector<uint8_t> rgb;
    std::vector<float> depth;
    std::vector<float> normal;
    std::vector<int32_t> object_id;
};

struct RayHits {  // include/neraulgraphics/renderer.h:88-95
    Memory memory;
    int device_id, count;
    const float *distance;
    const float *normal;
    const int32_t *object_id;
};

struct HostRayHits {  // include/neraulgraphics/renderer.h:97-101
    std::vector<float> distance;
    std::vector<float> normal;
    std::vector<int32_t> object_id;
};

class BatchRenderer {
public:
    // world_count worlds, each holding the instances (the first instances_in_world[w] of them when given),
    // cameras_per_world views each, on HIP device device_id.
    BatchRenderer(int width, int height, int world_count, int cameras_per_world, const std::vector<Sphere> &spheres,
                  const std::vector<Light> &lights, const Plane &floor, const std::vector<Mesh> &meshes,
                  const std::vector<Instance> &instances, const std::vector<int> &instances_in_world = {}, int device_id = 0)
        : instance_materials_(materials_of(instances)), sphere_materials_(materials_of(spheres)), floor_(floor), lights_(lights) {
        std::vector<float> verts, normals, sph, tf;
        std::vector<uint32_t> idx;
        std::vector<int32_t> vo{0}, io{0}, inst_mesh;
        std::vector<uint8_t> has;
        bool any_normals = false;
        for (const Mesh &m : meshes) any_normals = any_normals || !m.normals.empty();
        for (const Mesh &m : meshes) {
            if (m.vertices.size() % 3 != 0) throw std::invalid_argument("a mesh's vertices must be 3 floats each");
            if (!m.normals.empty() && m.normals.size() != m.vertices.size())
                throw std::invalid_argument("a mesh's normals must match its vertices");
            verts.insert(verts.end(), m.vertices.begin(), m.vertices.end());
            idx.insert(idx.end(), m.indices.begin(), m.indices.end());
            if (any_normals) {
                if (m.normals.empty()) normals.insert(normals.end(), m.vertices.size(), 0.0f);
                else normals.insert(normals.end(), m.normals.begin(), m.normals.end());
            }
            has.push_back(m.normals.empty() ? 0 : 1);
            vo.push_back((int32_t)(verts.size() / 3));
            io.push_back((int32_t)idx.size());
        }
        for (const Instance &i : instances) {
            inst_mesh.push_back(i.mesh);
            tf.insert(tf.end(), i.transform, i.transform + 12);
        }
        for (const Sphere &s : spheres) sph.insert(sph.end(), {s.center.x, s.center.y, s.center.z, s.radius});
        if (!instances_in_world.empty() && (int)instances_in_world.size() != world_count)
            throw std::invalid_argument("instances_in_world needs one count per world");
        ngx_scene_desc d{};
        d.width = width;
        d.height = height;
        d.worlds = world_count;
        d.cameras_per_world = cameras_per_world;
        d.mesh_count = (int32_t)meshes.size();
        d.vertices = verts.data();
        d.vertex_offsets = vo.data();
        d.indices = idx.data();
        d.index_offsets = io.data();
        d.normals = any_normals ? normals.data() : nullptr;
        d.mesh_has_normals = has.data();
        d.instances_per_world = (int32_t)instances.size();
        d.instance_mesh = inst_mesh.data();
        d.instance_transforms = tf.data();
        d.instances_in_world = instances_in_world.empty() ? nullptr : instances_in_world.data();
        d.sphere_count = (int32_t)spheres.size();
        d.spheres = sph.data();
        d.floor_y = floor.y;
        d.floor_half_extent_x = floor.half_extent_x;
        d.floor_z_min = floor.z_min;
        d.floor_z_max = floor.z_max;
        d.device_id = device_id;
        check(ngx_create(&d, &ctx_));
        try {
            apply_shading();
        } catch (...) {
            ngx_destroy(ctx_);
            throw;
        }
    }
    ~BatchRenderer() { ngx_destroy(ctx_); }
    BatchRenderer(BatchRenderer &&o) noexcept
        : ctx_(std::exchange(o.ctx_, nullptr)), instance_materials_(std::move(o.instance_materials_)),
          sphere_materials_(std::move(o.sphere_materials_)), floor_(o.floor_), lights_(std::move(o.lights_)), ambient_(o.ambient_),
          environment_(o.environment_) {}
    BatchRenderer &operator=(BatchRenderer &&o) noexcept {
        std::swap(ctx_, o.ctx_);
        std::swap(instance_materials_, o.instance_materials_);
        std::swap(sphere_materials_, o.sphere_materials_);
        std::swap(floor_, o.floor_);
        std::swap(lights_, o.lights_);
        std::swap(ambient_, o.ambient_);
        std::swap(environment_, o.environment_);
        return *this;
    }

    // The RGB picture's ambient level and sky (batch.h:206-212); refused values change nothing.
    void set_ambient(const float ambient) {
        const float old = ambient_;
        ambient_ = ambient;
        try {
            apply_shading();
        } catch (...) {
            ambient_ = old;
            throw;
        }
    }
    float ambient() const { return ambient_; }
    void set_environment(const Environment &environment) {
        const Environment old = environment_;
        environment_ = environment;
        try {
            apply_shading();
        } catch (...) {
            environment_ = old;
            throw;
        }
    }
    BatchRenderer(const BatchRenderer &) = delete;
    BatchRenderer &operator=(const BatchRenderer &) = delete;

    // world_count x instances_per_world x 12 floats, world-major (batch.h:173-179).
    void set_instance_transforms(const float *matrices, const Memory memory = Memory::Host) {
        check(ngx_set_instance_matrices(live(), matrices, mem(memory), nullptr));
        check(ngx_synchronize(ctx_));
    }
    void bind_bodies(const BodyBinding &b) {
        const size_t n = (size_t)instances_per_world();
        if (b.body_of_instance.size() != n || b.local_pose.size() != 7 * n || b.scale.size() != 3 * n)
            throw std::invalid_argument("a body binding needs one body index, 7 pose floats and 3 scale floats for each of the " +
                                        std::to_string(n) + " instances of a world");
        check(ngx_bind_bodies(live(), b.bodies_per_world, b.body_of_instance.data(), b.local_pose.data(), b.scale.data()));
    }
    // world_count x bodies_per_world x 7 floats: x, y, z, qx, qy, qz, qw.
    void set_body_poses(const float *poses, const Memory memory = Memory::Host) {
        check(ngx_set_body_poses(live(), poses, mem(memory), nullptr));
        check(ngx_synchronize(ctx_));
    }
    // Simulator geoms (the MuJoCo Warp flow): instance i of every world is drawn at geom geom_of_instance[i]'s world
    // frame, scaled by scale[3 i .. 3 i + 2] along the mesh's own axes. Independent of bind_bodies.
    void bind_geoms(const int geoms_per_world, const std::vector<int32_t> &geom_of_instance, const std::vector<float> &scale) {
        const size_t n = (size_t)instances_per_world();
        if (geom_of_instance.size() != n || scale.size() != 3 * n)
            throw std::invalid_argument("a geom binding needs one geom index and 3 scale floats for each of the " +
                                        std::to_string(n) + " instances of a world");
        check(ngx_bind_geoms(live(), geoms_per_world, geom_of_instance.data(), scale.data()));
    }
    // world_count x geoms_per_world x 3 positions and x 9 row-major rotations (MjData.geom_xpos, geom_xmat in FP32).
    void set_geom_frames(const float *xpos, const float *xmat, const Memory memory = Memory::Host) {
        check(ngx_set_geom_frames(live(), xpos, xmat, mem(memory), nullptr));
        check(ngx_synchronize(ctx_));
    }
    // No counterpart in the reference: moves mesh `mesh`'s vertices (as many as it was made with, 3 floats each) and refits
    // its bottom level on the device. Set the instance transforms or body poses again before the next render.
    void set_mesh_vertices(const int mesh, const float *vertices, const Memory memory = Memory::Host) {
        check(ngx_set_mesh_vertices(live(), mesh, vertices, mem(memory), nullptr));
        check(ngx_synchronize(ctx_));
    }
    // view_count x 12 camera-to-world floats; results stay in the renderer's buffers until the next render.
    BatchFrame render(const float *camera_matrices, const float fov, const int samples_per_pixel, const unsigned outputs,
                      const int frame = 0, const Memory camera_memory = Memory::Host) {
        ngx_batch_frame f{};
        check(ngx_render_batch(live(), camera_matrices, mem(camera_memory), fov, samples_per_pixel, outputs, frame, nullptr, &f));
        check(ngx_synchronize(ctx_));
        return BatchFrame{Memory::Device, f.device_id, f.views, f.width, f.height, f.rgb, f.depth, f.normal, f.object_id};
    }
    HostBatchFrame render_host(const float *camera_matrices, const float fov, const int samples_per_pixel, const unsigned outputs,
                               const int frame = 0) {
        const BatchFrame f = render(camera_matrices, fov, samples_per_pixel, outputs, frame);
        HostBatchFrame h;
        h.views = f.views;
        h.width = f.width;
        h.height = f.height;
        const size_t px = (size_t)f.views * f.width * f.height;
        if (f.rgb) h.rgb = copy_out(f.rgb, 3 * px);
        if (f.depth) h.depth = copy_out(f.depth, px);
        if (f.normal) h.normal = copy_out(f.normal, 3 * px);
        if (f.object_id) h.object_id = copy_out(f.object_id, px);
        return h;
    }
    // origins and directions: world_count x rays_per_world x 3 floats.
    RayHits cast_rays(const float *origins, const float *directions, const int rays_per_world, const float max_distance,
                      const Memory input_memory = Memory::Host) {
        ngx_ray_hits h{};
        check(ngx_cast_rays_batch(live(), origins, directions, rays_per_world, max_distance, mem(input_memory), nullptr, &h));
        check(ngx_synchronize(ctx_));
        return RayHits{Memory::Device, h.device_id, h.count, h.distance, h.normal, h.object_id};
    }
    HostRayHits cast_rays_host(const std::vector<float> &origins, const std::vector<float> &directions, const int rays_per_world,
                               const float max_distance) {
        if (origins.size() != (size_t)world_count() * rays_per_world * 3 || directions.size() != origins.size())
            throw std::invalid_argument("origins and directions must hold world_count x rays_per_world x 3 floats");
        const RayHits r = cast_rays(origins.data(), directions.data(), rays_per_world, max_distance);
        HostRayHits h;
        h.distance = copy_out(r.distance, (size_t)r.count);
        h.normal = copy_out(r.normal, 3 * (size_t)r.count);
        h.object_id = copy_out(r.object_id, (size_t)r.count);
        return h;
    }

    int width() const { return frame_size().first; }
    int height() const { return frame_size().second; }
    int world_count() const { return ngx_world_count(ctx_); }
    int view_count() const { return ngx_view_count(ctx_); }
    int cameras_per_world() const { return view_count() / world_count(); }
    int instances_per_world() const { return ngx_instances_per_world(ctx_); }
    Memory memory() const { return Memory::Device; }
    float last_kernel_ms() const { return ngx_last_kernel_ms(ctx_); }
    ngx_ctx *handle() const { return ctx_; }

private:
    ngx_ctx *ctx_ = nullptr;
    std::vector<ngx_material> instance_materials_, sphere_materials_;
    Plane floor_;
    std::vector<Light> lights_;
    float ambient_ = 0.0f;
    Environment environment_{};

    static ngx_material to_c(const Material &m) {
        return ngx_material{{m.diffuse_color.x, m.diffuse_color.y, m.diffuse_color.z}, m.diffuse_albedo, m.specular_albedo, m.reflectivity,
                            m.specular_exponent};
    }
    template <class T>
    static std::vector<ngx_material> materials_of(const std::vector<T> &items) {
        std::vector<ngx_material> out;
        for (const T &i : items) out.push_back(to_c(i.material));
        return out;
    }
    void apply_shading() {
        std::vector<ngx_light> ls;
        for (const Light &l : lights_)
            ls.push_back(ngx_light{{l.position.x, l.position.y, l.position.z}, l.intensity, l.radius, l.falloff,
                                   {l.direction.x, l.direction.y, l.direction.z}, l.spot_angle, l.spot_softness,
                                   {l.color.x, l.color.y, l.color.z}});
        ngx_shading h{};
        h.instance_materials = instance_materials_.data();
        h.sphere_materials = sphere_materials_.data();
        h.floor_material = to_c(floor_.material);
        h.floor_checker_scale = floor_.checker_scale;
        h.light_count = (int32_t)ls.size();
        h.lights = ls.data();
        h.ambient = ambient_;
        const Environment &e = environment_;
        const float sky[10] = {e.up.x, e.up.y, e.up.z, e.color_up.x, e.color_up.y, e.color_up.z, e.color_down.x, e.color_down.y,
                               e.color_down.z, e.intensity};
        std::copy(sky, sky + 3, h.sky_up);
        std::copy(sky + 3, sky + 6, h.sky_color_up);
        std::copy(sky + 6, sky + 9, h.sky_color_down);
        h.sky_intensity = sky[9];
        check(ngx_set_shading(live(), &h));
    }

    ngx_ctx *live() const {
        if (!ctx_) throw std::logic_error("this BatchRenderer was moved from");
        return ctx_;
    }
    static ngx_memory mem(const Memory m) { return m == Memory::Host ? NGX_MEM_HOST : NGX_MEM_DEVICE; }
    template <class T>
    std::vector<T> copy_out(const T *src, const size_t n) const {
        std::vector<T> v(n);
        if (n) check(ngx_copy_to_host(ctx_, src, v.data(), n * sizeof(T)));
        return v;
    }
    std::pair<int, int> frame_size() const {
        int32_t w = 0, h = 0;
        check(ngx_frame_size(ctx_, &w, &h));
        return {w, h};
    }
};

}  // namespace ngx

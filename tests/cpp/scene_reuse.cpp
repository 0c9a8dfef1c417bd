// Scenes built one after another at the same address (a scene_t on the
// stack in a loop), each with other geometry: every render must use its own
// scene's meshes, not a device copy left by an earlier scene object.  The
// last two renders edit one scene's arrays: first by replacing its index
// array (noticed without a generation bump), then by editing vertex values
// in place and bumping device_generation.
//
// usage: scene_reuse <width> <height>
// prints one line per render: the number of pixels with a depth written
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <swegl/swegl.hpp>

static swegl::primitive_t quad(float half)
{
    swegl::primitive_t p;
    const float xs[4] = {-half, half, -half, half}, ys[4] = {-half, -half, half, half};
    for (int i = 0; i < 4; i++) {
        swegl::mesh_vertex_t v{};
        v.v = swegl::vertex_t(xs[i], ys[i], 0.0f);
        v.normal = swegl::normal_t(0, 0, 1);
        v.tex_coords = swegl::vec2f_t(0.5f, 0.5f);
        p.vertices.push_back(v);
    }
    p.indices = {0, 1, 2, 2, 1, 3};
    p.mode = swegl::primitive_t::TRIANGLES;
    p.material_id = 0;
    return p;
}

static long covered(const swegl::viewport_t &vp, int w, int h)
{
    long n = 0;
    const float *z = vp.zbuffer();
    for (int i = 0; i < w * h; i++)
        n += z[i] < 1e30f;
    return n;
}

int main(int argc, char **argv)
{
    if (argc != 3)
        return 2;
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]);
    swegl::owned_surface surf(w, h);
    std::shared_ptr<swegl::pixel_shader_t> shader = std::make_shared<swegl::pixel_shader_t>();
    swegl::viewport_t viewport(0, 0, w, h, &surf.surface, shader, 0);
    swegl::post_shader_t post_null;
    viewport.set_post_shader(post_null);
    viewport.m_camera.translate(0, 0, -3);
    const float halves[3] = {0.25f, 0.5f, 1.0f};
    try {
        for (int k = 0; k < 3; k++) {
            swegl::scene_t scene;   // same stack slot every iteration
            scene.materials.push_back(swegl::material_t{swegl::pixel_colors(200, 100, 50, 255), 1, 1, -1, false});
            swegl::node_t n;
            n.primitives.push_back(quad(halves[k]));
            scene.nodes.push_back(n);
            scene.root_nodes.push_back(0);
            scene.ambient_light_intensity = 1.0f;
            swegl::render(scene, viewport);
            std::printf("%ld\n", covered(viewport, w, h));
            if (k == 2) {
                // a replaced index array (one triangle of the quad): no bump needed
                scene.nodes[0].primitives[0].indices = std::vector<swegl::vertex_idx>{0, 1, 2};
                swegl::render(scene, viewport);
                std::printf("%ld\n", covered(viewport, w, h));
                // vertex values edited in place: the generation bump makes it visible
                for (auto &v : scene.nodes[0].primitives[0].vertices)
                    v.v = swegl::vertex_t(v.v.x() * 0.5f, v.v.y() * 0.5f, 0.0f);
                scene.device_generation++;
                swegl::render(scene, viewport);
                std::printf("%ld\n", covered(viewport, w, h));
            }
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "render failed: %s\n", e.what());
        return 1;
    }
    return 0;
}

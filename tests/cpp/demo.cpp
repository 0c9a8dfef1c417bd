// The reference demo's frame loop without a window (src/test_1.cpp:50-148,
// 324-387): its scene (torus, stretched cube, sphere, three overlapping
// triangles, two of them translucent, two small cubes at the point
// lights), an 800x600 viewport with 3 transparency layers, Phong + bilinear
// shading; each frame renders into the host surface (the SDL surface of the
// demo), animates the scene and turns the camera as a mouse move would.
// Prints one JSON line with the frame rate; writes the last frame's pixels
// and depth as raw files.
//
// usage: demo <tex0.raw> <tex1.raw> <tex2.raw> (256x256 u32 each) <frames> <width> <height>
//             <out_pixels.raw> <out_z.raw>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <swegl/swegl.hpp>

static std::vector<unsigned int> read_raw(const char *path)
{
    std::vector<unsigned int> t(256 * 256);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(t.data(), 4, t.size(), f) != t.size())
        std::exit(3);
    std::fclose(f);
    return t;
}

static swegl::scene_t build_scene(const char *const *tex)
{
    swegl::scene_t s;
    for (int k = 0; k < 3; k++) {
        std::vector<unsigned int> t = read_raw(tex[k]);
        s.images.emplace_back(t.data(), 256, 256);
    }
    const int cols[8][5] = {{128, 128, 128, 255, 0}, {128, 128, 128, 255, 1}, {128, 128, 128, 255, 2},
                            {128, 128, 255, 255, -1}, {255, 128, 255, 255, -1}, {128, 128, 255, 100, -1},
                            {128, 255, 128, 100, -1}, {255, 128, 128, 100, -1}};
    for (const auto &c : cols)
        s.materials.push_back(swegl::material_t{
            swegl::pixel_colors((unsigned char)c[0], (unsigned char)c[1], (unsigned char)c[2], (unsigned char)c[3]),
            1, 1, c[4]});
    s.ambient_light_intensity = 0.2f;
    s.sun_direction = swegl::normal_t{1.0, -1.0, -1.0};
    s.sun_direction.normalize();
    s.sun_intensity = 0.3f;
    s.point_source_lights.push_back({swegl::vertex_t(0.0f, 3.0f, 0.0f), 0.6f});
    s.point_source_lights.push_back({swegl::vertex_t(0.5f, 2.0f, 0.0f), 100.0f});

    swegl::node_t tore = swegl::make_tore(100, 1);
    tore.rotation = swegl::matrix44_t::identity();
    tore.rotation.rotate_z(0.5f);
    tore.translation = swegl::vertex_t(0.0f, 0.0f, -2.5f);
    s.nodes.push_back(std::move(tore));
    swegl::node_t cube = swegl::make_cube(1.0f, 0);
    cube.scale.x() = 2;
    s.nodes.push_back(std::move(cube));
    swegl::node_t sphere = swegl::make_sphere(100, 2.0f, 2);
    sphere.translation = swegl::vertex_t(3.0f, 0.0f, -1.0f);
    s.nodes.push_back(std::move(sphere));
    swegl::node_t tri = swegl::make_tri(1, 5);
    tri.translation = swegl::vertex_t(1.0f, 0.5f, 2.1f);
    s.nodes.push_back(std::move(tri));
    swegl::node_t cube2 = swegl::make_cube(0.1f, 4);
    cube2.translation = s.point_source_lights[0].position;
    s.nodes.push_back(std::move(cube2));
    swegl::node_t cube3 = swegl::make_cube(0.1f, 4);
    cube3.translation = s.point_source_lights[1].position;
    s.nodes.push_back(std::move(cube3));
    swegl::node_t tri2 = swegl::make_tri(1, 6);
    tri2.translation = swegl::vertex_t(1.0f, 0.5f, 2.0f);
    s.nodes.push_back(std::move(tri2));
    swegl::node_t tri3 = swegl::make_tri(1, 7);
    tri3.translation = swegl::vertex_t(1.0f, 0.5f, 2.2f);
    s.nodes.push_back(std::move(tri3));
    for (int i = 0; i < (int)s.nodes.size(); i++)
        s.root_nodes.push_back(i);
    return s;
}

int main(int argc, char **argv)
{
    if (argc != 9) {
        std::fprintf(stderr, "usage: %s tex0 tex1 tex2 frames w h out_pixels out_z\n", argv[0]);
        return 2;
    }
    const int frames = std::atoi(argv[4]), w = std::atoi(argv[5]), h = std::atoi(argv[6]);
    try {
        swegl::scene_t scene = build_scene(argv + 1);
        swegl::owned_surface surf(w, h);
        std::shared_ptr<swegl::pixel_shader_t> shader = std::make_shared<
            swegl::pixel_shader_light_and_texture<swegl::pixel_shader_lights_phong,
                                                  swegl::pixel_shader_texture_bilinear>>();
        swegl::viewport_t viewport(0, 0, w, h, &surf.surface, shader, 3);
        swegl::post_shader_t post_null;
        viewport.set_post_shader(post_null);
        viewport.m_camera.translate(1, 2, -5);
        viewport.m_camera.rotate_y(-0.2f);
        viewport.m_camera.rotate_x(-0.3f);

        swegl::render(scene, viewport);   // first frame: uploads the scene
        const auto t0 = std::chrono::steady_clock::now();
        for (int f = 1; f < frames; f++) {
            viewport.m_camera.rotate_y(0.001f);   // a small mouse move per frame
            swegl::render(scene, viewport);
            scene.animate(f / 60.0f);
        }
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("{\"frames\": %d, \"seconds\": %.6f, \"fps\": %.2f, \"width\": %d, \"height\": %d}\n",
                    frames - 1, secs, frames > 1 ? (frames - 1) / secs : 0.0, w, h);
        FILE *o = std::fopen(argv[7], "wb");
        std::fwrite(surf.storage.data(), 4, surf.storage.size(), o);
        std::fclose(o);
        FILE *z = std::fopen(argv[8], "wb");
        std::fwrite(viewport.zbuffer(), 4, (size_t)w * h, z);
        std::fclose(z);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "demo: %s\n", e.what());
        return 1;
    }
    return 0;
}

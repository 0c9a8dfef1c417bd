// A C++ program that loads a glTF/GLB scene with swegl::load_scene, poses it
// with scene_t::animate and renders it with swegl::render, as the
// reference demo does with a file given on its command line
// (src/test_1.cpp:330-374).  Lights are fixed here (the loader leaves them
// to the caller); outputs are raw files for a test to compare.
//
// usage: render_asset <scene.glb|.gltf> <seconds> <width> <height> <out_pixels.raw> <out_z.raw>
#include <cstdio>
#include <cstdlib>

#include <swegl/swegl.hpp>

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s scene seconds w h out_pixels out_z\n", argv[0]);
        return 2;
    }
    const int w = std::atoi(argv[3]), h = std::atoi(argv[4]);
    try {
        swegl::scene_t scene = swegl::load_scene(argv[1]);
        scene.ambient_light_intensity = 0.15f;
        scene.sun_direction = swegl::normal_t(0.3f, -0.8f, -0.4f);
        scene.sun_intensity = 0.9f;
        scene.point_source_lights.push_back({swegl::vertex_t(0.5f, 1.0f, -1.5f), 1.5f});
        scene.animate((float)std::atof(argv[2]));

        swegl::owned_surface surf(w, h);
        std::shared_ptr<swegl::pixel_shader_t> shader = std::make_shared<
            swegl::pixel_shader_light_and_texture<swegl::pixel_shader_lights_phong,
                                                  swegl::pixel_shader_texture_bilinear>>();
        swegl::viewport_t viewport(0, 0, w, h, &surf.surface, shader, 0);
        swegl::post_shader_t post_null;
        viewport.set_post_shader(post_null);
        viewport.m_camera.translate(0, 0.3f, -4);
        swegl::render(scene, viewport);

        FILE *o = std::fopen(argv[5], "wb");
        std::fwrite(surf.storage.data(), 4, surf.storage.size(), o);
        std::fclose(o);
        FILE *z = std::fopen(argv[6], "wb");
        std::fwrite(viewport.zbuffer(), 4, (size_t)w * h, z);
        std::fclose(z);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "render_asset: %s\n", e.what());
        return 1;
    }
    return 0;
}

// A C++ program using the swegl API the way the reference demo does
// (src/test_1.cpp:346-374): build a scene, make a viewport on a surface,
// set shaders and camera, call swegl::render.  Inputs and outputs are raw
// files so a test can compare the result with the oracle.
//
// usage: render_cube <texture.raw 256x256 u32> <rot_x> <rot_y> <rot_z>
//                    <sun_x> <sun_y> <sun_z> <width> <height> <out_pixels.raw> <out_z.raw>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <swegl/swegl.hpp>

int main(int argc, char **argv)
{
    if (argc != 12) {
        std::fprintf(stderr, "usage: %s tex rx ry rz sx sy sz w h out_pixels out_z\n", argv[0]);
        return 2;
    }
    std::vector<unsigned int> tex(256 * 256);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(tex.data(), 4, tex.size(), f) != tex.size())
        return 3;
    std::fclose(f);
    const int w = std::atoi(argv[8]), h = std::atoi(argv[9]);

    swegl::scene_t scene;
    scene.images.emplace_back(tex.data(), 256, 256);
    scene.materials.push_back(swegl::material_t{swegl::pixel_colors(255, 255, 255, 255), 1, 1, 0, false});
    swegl::node_t cube = swegl::make_cube(1.0f, 0);
    cube.rotation = swegl::matrix44_t::identity();
    cube.rotation.rotate_x((float)std::atof(argv[2]));
    cube.rotation.rotate_y((float)std::atof(argv[3]));
    cube.rotation.rotate_z((float)std::atof(argv[4]));
    scene.nodes.push_back(cube);
    scene.root_nodes.push_back(0);
    scene.ambient_light_intensity = 0.1f;
    scene.sun_direction = swegl::normal_t((float)std::atof(argv[5]), (float)std::atof(argv[6]),
                                          (float)std::atof(argv[7]));
    scene.sun_intensity = 1.0f;

    swegl::owned_surface surf(w, h);
    std::shared_ptr<swegl::pixel_shader_t> shader = std::make_shared<
        swegl::pixel_shader_light_and_texture<swegl::pixel_shader_lights_phong, swegl::pixel_shader_texture_bilinear>>();
    swegl::viewport_t viewport(0, 0, w, h, &surf.surface, shader, 0);
    swegl::post_shader_t post_null;
    viewport.set_post_shader(post_null);
    viewport.m_camera.translate(0, 0, -3);

    try {
        swegl::render(scene, viewport);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "render failed: %s\n", e.what());
        return 1;
    }
    FILE *o = std::fopen(argv[10], "wb");
    std::fwrite(surf.storage.data(), 4, surf.storage.size(), o);
    std::fclose(o);
    FILE *z = std::fopen(argv[11], "wb");
    std::fwrite(viewport.zbuffer(), 4, (size_t)w * h, z);
    std::fclose(z);
    return 0;
}

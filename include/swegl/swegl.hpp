// swegl/swegl.hpp -- the swegl API on MI355X (umbrella header, as the
// reference's swegl/swegl.hpp).  Link with -lswegl_hip.
#pragma once

#include "data/gltf.hpp"
#include "data/model.hpp"
#include "misc/image.hpp"
#include "projection/camera.hpp"
#include "projection/points.hpp"
#include "render/colors.hpp"
#include "render/renderer.hpp"
#include "render/surface.hpp"
#include "render/viewport.hpp"

// asset.cpp -- glTF 2.0 / GLB scenes into host arrays ready for
// swegl_scene_upload (reference: src/data/gltf.cpp, swegl/data/gltf.hpp),
// node animation (scene_t::animate, swegl/data/model.hpp:146-177) and the
// texture readers behind them (src/misc/image.cpp).
//
// The loader keeps the reference's reading of a file, including the parts
// that differ from the glTF specification, because a drop-in must build the
// same scene from the same file:
//  * TEXCOORD_0 is read transposed: tex_coords = (float at +4, float at +0)
//    (gltf.cpp:203-206);
//  * indices are read as 16-bit words at the accessor's stride whatever the
//    component type (gltf.cpp:214-218): 32-bit indices keep their low half;
//  * every primitive gets two extra zero vertices past the accessor's count
//    (gltf.cpp:178-179, room for the near-clip vertices);
//  * a rotation quaternion (x, y, z, w) is passed to from_quaternion as
//    (q0, q1, q2, q3) with q0 taken as the scalar part (gltf.cpp:231-239,
//    matrix44.hpp:34-43);
//  * a node matrix is split into translation, the lengths of its first three
//    rows as scale, and the columns divided by them as rotation
//    (gltf.cpp:251-305);
//  * a material without pbrMetallicRoughness is white and one-sided even
//    when doubleSided is set (gltf.cpp:117-124);
//  * a mesh without primitives is skipped, so later mesh indices shift, and
//    a mesh shared by several nodes goes to the first of them only (its
//    primitives are moved, gltf.cpp:153-157, 307-308).
// Where the reference's behaviour is undefined (reads outside a buffer,
// indices past the vertices, missing required members) the loader returns
// SWEGL_EINVAL with a message instead.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/swegl_hip.h"
#include "image.h"
#include "json.h"

int sw_fail_msg(int rc, const std::string &msg);

struct swegl_asset {
    std::vector<swegl_node_desc> nodes;
    std::vector<int32_t> roots, children;
    std::vector<swegl_primitive_desc> prims;
    std::vector<std::vector<float>> pos, nor, uv;
    std::vector<std::vector<uint32_t>> idx;
    std::vector<swegl_material_desc> materials;
    std::vector<std::vector<uint32_t>> bitmaps;
    std::vector<swegl_texture_desc> textures;
    std::vector<std::vector<swegl_anim_step>> steps;
    std::vector<std::vector<swegl_anim_channel>> channels;
    std::vector<swegl_animation> animations;
};

namespace {

struct Error {
    int rc;
    std::string msg;
};
[[noreturn]] void bad(const std::string &m, int rc = SWEGL_EINVAL) { throw Error{rc, m}; }

std::vector<uint8_t> read_file(const std::string &path)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (!f)
        bad("cannot open " + path);
    std::vector<uint8_t> d;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0)
        d.insert(d.end(), buf, buf + n);
    fclose(f);
    return d;
}

std::string lower(std::string s)
{
    for (char &c : s)
        c = (char)tolower((unsigned char)c);
    return s;
}
bool ends_with(const std::string &s, const char *t)
{
    size_t n = strlen(t);
    return s.size() >= n && s.compare(s.size() - n, n, t) == 0;
}
std::string parent_dir(const std::string &p)
{
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

const sj::Value &need(const sj::Value &o, const char *key, const char *where)
{
    const sj::Value *v = o.get(key);
    if (!v)
        bad(std::string(where) + " has no \"" + key + "\"");
    return *v;
}
int as_int(const sj::Value &v, const char *what)
{
    if (!v.is_num() || !(v.num > -2147483649.0 && v.num < 2147483648.0))
        bad(std::string(what) + " is not an integer");
    return (int)v.num;
}
float as_float(const sj::Value &v, const char *what)
{
    if (!v.is_num())
        bad(std::string(what) + " is not a number");
    return (float)v.num;
}
int int_or(const sj::Value &o, const char *key, int dflt)
{
    const sj::Value *v = o.get(key);
    return v ? as_int(*v, key) : dflt;
}
// element i of an array member, range-checked
const sj::Value &at(const sj::Value &arr, long long i, const char *what)
{
    if (!arr.is_arr() || i < 0 || (size_t)i >= arr.arr.size())
        bad(std::string(what) + " index " + std::to_string(i) + " out of range");
    return arr.arr[(size_t)i];
}

// x86 (unsigned char)(float): cvttss2si, then the low byte
uint8_t to_u8(float x)
{
    int32_t i = (x > -2147483649.0f && x < 2147483648.0f) ? (int32_t)x : INT32_MIN;
    return (uint8_t)i;
}

struct Bytes {
    const uint8_t *p = nullptr;
    size_t n = 0;   // bytes readable from p (to the end of the underlying buffer)
    size_t len = 0; // bytes of the view itself
};

struct Accessor {
    Bytes data;      // the accessor's view: starts at its byteOffset
    std::string type;
    int component = 0, count = 0, stride = 0;
};

int component_size(int c)
{
    switch (c) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: bad("accessor componentType " + std::to_string(c) + " unknown");
    }
}
int type_count(const std::string &t)
{
    if (t == "SCALAR") return 1;
    if (t == "VEC2") return 2;
    if (t == "VEC3") return 3;
    if (t == "VEC4") return 4;
    if (t == "MAT2") return 4;
    if (t == "MAT3") return 9;
    if (t == "MAT4") return 16;
    bad("accessor type " + t + " unknown");
}

// bytes [off, off+size) of an accessor, checked against the buffer
const uint8_t *read_at(const Accessor &a, size_t off, size_t size)
{
    if (off + size > a.data.n)
        bad("accessor read past the end of its buffer");
    return a.data.p + off;
}
float f32_at(const Accessor &a, size_t off)
{
    float v;
    memcpy(&v, read_at(a, off, 4), 4);
    return v;
}

// matrix44_t::from_quaternion (matrix44.hpp:34-43)
void from_quaternion(float q0, float q1, float q2, float q3, float *m)
{
    const float r[16] = {2 * (q0 * q0 + q1 * q1) - 1, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2), 0.0f,
                         2 * (q1 * q2 + q0 * q3), 2 * (q0 * q0 + q2 * q2) - 1, 2 * (q2 * q3 - q0 * q1), 0.0f,
                         2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), 2 * (q0 * q0 + q3 * q3) - 1, 0.0f,
                         0.0f, 0.0f, 0.0f, 1.0f};
    memcpy(m, r, sizeof r);
}

// vector_t::len (points.hpp:71-74)
float len3(float x, float y, float z) { return (float)std::sqrt(x * x + y * y + z * z); }

void identity(float *m)
{
    static const float I[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    memcpy(m, I, sizeof I);
}

struct Loader {
    std::string filename, root;
    sj::Value j;
    std::vector<std::vector<uint8_t>> files;   // external buffers and the GLB itself
    std::vector<Bytes> buffers;
    std::vector<Bytes> views;
    std::vector<int> view_stride;
    std::vector<Accessor> accessors;
    swegl_asset &A;

    explicit Loader(swegl_asset &a) : A(a) {}

    void load_buffers(const Bytes *glb_bin)
    {
        const sj::Value *jb = j.get("buffers");
        if (!jb)
            return;
        for (const sj::Value &b : jb->arr) {
            const int len = as_int(need(b, "byteLength", "buffer"), "byteLength");
            if (len < 0)
                bad("negative buffer byteLength");
            if (const sj::Value *uri = b.get("uri")) {
                if (!uri->is_str())
                    bad("buffer uri is not a string");
                if (uri->str.compare(0, 5, "data:") == 0)
                    bad("data: URIs are not read (the reference reads buffers from files)", SWEGL_ENOTSUP);
                files.push_back(read_file(root + uri->str));
                if ((size_t)len > files.back().size())
                    bad("buffer " + uri->str + " is shorter than its byteLength");
                buffers.push_back(Bytes{files.back().data(), (size_t)len, (size_t)len});
            } else {
                if (!glb_bin)
                    bad("a buffer without uri outside a GLB file");
                if ((size_t)len > glb_bin->n)
                    bad("GLB BIN chunk is shorter than the buffer's byteLength");
                buffers.push_back(Bytes{glb_bin->p, (size_t)len, (size_t)len});
            }
        }
    }

    void load_views()
    {
        const sj::Value *jv = j.get("bufferViews");
        if (!jv)
            return;
        for (const sj::Value &v : jv->arr) {
            const int b = as_int(need(v, "buffer", "bufferView"), "buffer");
            if (b < 0 || (size_t)b >= buffers.size())
                bad("bufferView buffer out of range");
            const int off = int_or(v, "byteOffset", 0), len = as_int(need(v, "byteLength", "bufferView"), "byteLength");
            if (off < 0 || len < 0 || (size_t)off > buffers[b].n)
                bad("bufferView outside its buffer");
            views.push_back(Bytes{buffers[b].p + off, buffers[b].n - off, (size_t)len});
            view_stride.push_back(int_or(v, "byteStride", 0));
        }
    }

    // gltf.cpp:13-52
    void load_accessors()
    {
        const sj::Value *ja = j.get("accessors");
        if (!ja)
            return;
        for (const sj::Value &a : ja->arr) {
            Accessor acc;
            const sj::Value &t = need(a, "type", "accessor");
            if (!t.is_str())
                bad("accessor type is not a string");
            acc.type = t.str;
            acc.component = as_int(need(a, "componentType", "accessor"), "componentType");
            acc.count = as_int(need(a, "count", "accessor"), "count");
            const sj::Value *bv = a.get("bufferView");
            if (!bv)
                bad("accessors without bufferView (sparse or zero-filled) are not read", SWEGL_ENOTSUP);
            const int vi = as_int(*bv, "bufferView");
            if (vi < 0 || (size_t)vi >= views.size())
                bad("accessor bufferView out of range");
            acc.stride = view_stride[vi];
            if (acc.stride == 0)
                acc.stride = component_size(acc.component) * type_count(acc.type);
            const int off = int_or(a, "byteOffset", 0);
            if (off < 0 || (size_t)off > views[vi].n || acc.count < 0)
                bad("accessor outside its bufferView");
            acc.data = Bytes{views[vi].p + off, views[vi].n - off, (size_t)acc.stride * acc.count};
            accessors.push_back(acc);
        }
    }

    const Accessor &accessor(const sj::Value &v, const char *what)
    {
        const int i = as_int(v, what);
        if (i < 0 || (size_t)i >= accessors.size())
            bad(std::string(what) + " accessor out of range");
        return accessors[i];
    }

    void add_image(const std::vector<uint8_t> &bytes, bool png, const std::string &name)
    {
        std::vector<uint32_t> px;
        int w = 0, h = 0;
        std::string err;
        const bool ok = png ? sw_img::decode_png(bytes.data(), bytes.size(), px, w, h, err)
                            : sw_img::decode_bmp(bytes.data(), bytes.size(), px, w, h, err);
        if (!ok)
            bad("image " + name + ": " + err);
        A.bitmaps.push_back(std::move(px));
        A.textures.push_back(swegl_texture_desc{nullptr, (uint32_t)w, (uint32_t)h});
    }

    // gltf.cpp:73-111
    void load_images()
    {
        const sj::Value *ji = j.get("images");
        if (!ji)
            return;
        for (const sj::Value &im : ji->arr) {
            if (const sj::Value *uri = im.get("uri")) {
                if (!uri->is_str())
                    bad("image uri is not a string");
                const std::string l = lower(uri->str);
                if (ends_with(l, "png"))
                    add_image(read_file(root + uri->str), true, uri->str);
                else if (ends_with(l, "bmp"))
                    add_image(read_file(root + uri->str), false, uri->str);
                else if (ends_with(l, "jpg") || ends_with(l, "jpeg"))
                    bad("JPEG textures need a JPEG decoder, which this build does not have", SWEGL_ENOTSUP);
                else
                    bad("image format of " + uri->str + " is not read by the reference", SWEGL_ENOTSUP);
                continue;
            }
            const sj::Value *mime = im.get("mimeType");
            if (!mime || !mime->is_str())
                bad("image without uri or mimeType");
            const int vi = as_int(need(im, "bufferView", "image"), "bufferView");
            if (vi < 0 || (size_t)vi >= views.size())
                bad("image bufferView out of range");
            const Bytes &v = views[vi];
            if (v.len > v.n)
                bad("image bufferView outside its buffer");
            if (mime->str == "image/png")
                add_image(std::vector<uint8_t>(v.p, v.p + v.len), true, "bufferView " + std::to_string(vi));
            else if (mime->str == "image/jpg" || mime->str == "image/jpeg")
                bad("JPEG textures need a JPEG decoder, which this build does not have", SWEGL_ENOTSUP);
            else
                bad("image mimeType " + mime->str + " is not read by the reference", SWEGL_ENOTSUP);
        }
    }

    // gltf.cpp:113-148
    void load_materials()
    {
        const sj::Value *jm = j.get("materials");
        if (!jm)
            return;
        for (const sj::Value &m : jm->arr) {
            swegl_material_desc d{0xFFFFFFFFu, 1.0f, 1.0f, -1, 0};
            const sj::Value *ds = m.get("doubleSided");
            const bool double_sided = ds && ds->kind == sj::Value::Bool && ds->b;
            const sj::Value *pbr = m.get("pbrMetallicRoughness");
            if (!pbr) {
                A.materials.push_back(d);   // one-sided whatever doubleSided says
                continue;
            }
            if (const sj::Value *f = pbr->get("baseColorFactor")) {
                float c[4];
                for (int k = 0; k < 4; k++)
                    c[k] = as_float(at(*f, k, "baseColorFactor"), "baseColorFactor");
                d.color = (uint32_t)to_u8(255 * c[2]) | (uint32_t)to_u8(255 * c[1]) << 8 |
                          (uint32_t)to_u8(255 * c[0]) << 16 | (uint32_t)to_u8(255 * c[3]) << 24;
            }
            if (const sj::Value *t = pbr->get("baseColorTexture")) {
                const int ti = as_int(need(*t, "index", "baseColorTexture"), "texture index");
                const sj::Value &tex = at(need(j, "textures", "document"), ti, "textures");
                d.texture_idx = as_int(need(tex, "source", "texture"), "texture source");
            }
            if (const sj::Value *f = pbr->get("metallicFactor"))
                d.metallic = as_float(*f, "metallicFactor");
            if (const sj::Value *f = pbr->get("roughnessFactor"))
                d.roughness = as_float(*f, "roughnessFactor");
            d.double_sided = double_sided ? 1 : 0;
            A.materials.push_back(d);
        }
    }

    struct Prim {
        std::vector<float> pos, nor, uv;
        std::vector<uint32_t> idx;
        int mode = 4, material = -1;
    };

    // gltf.cpp:150-222
    std::vector<std::vector<Prim>> load_meshes()
    {
        std::vector<std::vector<Prim>> meshes;
        const sj::Value *jm = j.get("meshes");
        if (!jm)
            return meshes;
        for (const sj::Value &mesh : jm->arr) {
            const sj::Value *jp = mesh.get("primitives");
            if (!jp)
                continue;   // skipped: later mesh indices shift (gltf.cpp:153-154)
            meshes.emplace_back();
            for (const sj::Value &p : jp->arr) {
                Prim P;
                P.material = int_or(p, "material", -1);
                P.mode = int_or(p, "mode", 4);
                const sj::Value &attr = need(p, "attributes", "primitive");
                const Accessor &va = accessor(need(attr, "POSITION", "attributes"), "POSITION");
                const size_t nv = (size_t)va.count + 2;
                P.pos.assign(nv * 3, 0.0f);
                P.nor.assign(nv * 3, 0.0f);
                P.uv.assign(nv * 2, 0.0f);
                for (int i = 0; i < va.count; i++)
                    for (int k = 0; k < 3; k++)
                        P.pos[3 * i + k] = f32_at(va, (size_t)i * va.stride + 4 * k);
                if (const sj::Value *n = attr.get("NORMAL")) {
                    const Accessor &na = accessor(*n, "NORMAL");
                    if ((size_t)na.count > nv)
                        bad("more normals than vertices");
                    for (int i = 0; i < na.count; i++)
                        for (int k = 0; k < 3; k++)
                            P.nor[3 * i + k] = f32_at(na, (size_t)i * na.stride + 4 * k);
                }
                if (const sj::Value *t = attr.get("TEXCOORD_0")) {
                    const Accessor &ta = accessor(*t, "TEXCOORD_0");
                    if ((size_t)ta.count > nv)
                        bad("more texture coordinates than vertices");
                    for (int i = 0; i < ta.count; i++) {
                        P.uv[2 * i] = f32_at(ta, (size_t)i * ta.stride + 4);   // transposed as read
                        P.uv[2 * i + 1] = f32_at(ta, (size_t)i * ta.stride);
                    }
                }
                if (const sj::Value *ix = p.get("indices")) {
                    const Accessor &ia = accessor(*ix, "indices");
                    P.idx.resize(ia.count);
                    for (int i = 0; i < ia.count; i++) {
                        uint16_t v;
                        memcpy(&v, read_at(ia, (size_t)i * ia.stride, 2), 2);
                        P.idx[i] = v;
                    }
                }
                meshes.back().push_back(std::move(P));
            }
        }
        return meshes;
    }

    // gltf.cpp:224-318
    void load_nodes(std::vector<std::vector<Prim>> &meshes)
    {
        const sj::Value *jn = j.get("nodes");
        if (!jn)
            return;
        std::vector<std::vector<int32_t>> kids;
        std::vector<std::vector<Prim>> node_prims;
        for (const sj::Value &n : jn->arr) {
            swegl_node_desc d;
            memset(&d, 0, sizeof d);
            d.scale[0] = d.scale[1] = d.scale[2] = 1.0f;
            identity(d.rotation);
            if (const sj::Value *r = n.get("rotation")) {
                float q[4];
                for (int k = 0; k < 4; k++)
                    q[k] = as_float(at(*r, k, "rotation"), "rotation");
                from_quaternion(q[0], q[1], q[2], q[3], d.rotation);
            }
            if (const sj::Value *t = n.get("translation"))
                for (int k = 0; k < 3; k++)
                    d.translation[k] = as_float(at(*t, k, "translation"), "translation");
            if (const sj::Value *s = n.get("scale"))
                for (int k = 0; k < 3; k++)
                    d.scale[k] = as_float(at(*s, k, "scale"), "scale");
            if (const sj::Value *mj = n.get("matrix")) {
                float c[16], m[16];   // column-major in the file, row-major here
                for (int k = 0; k < 16; k++)
                    c[k] = as_float(at(*mj, k, "matrix"), "matrix");
                for (int r = 0; r < 4; r++)
                    for (int col = 0; col < 4; col++)
                        m[r * 4 + col] = c[col * 4 + r];
                d.translation[0] = m[3];
                d.translation[1] = m[7];
                d.translation[2] = m[11];
                d.scale[0] = len3(m[0], m[1], m[2]);
                d.scale[1] = len3(m[4], m[5], m[6]);
                d.scale[2] = len3(m[8], m[9], m[10]);
                m[3] = m[7] = m[11] = 0;
                for (int col = 0; col < 3; col++)
                    if (d.scale[col] != 0)
                        for (int r = 0; r < 3; r++)
                            m[r * 4 + col] /= d.scale[col];
                memcpy(d.rotation, m, sizeof m);
            }
            node_prims.emplace_back();
            if (const sj::Value *mi = n.get("mesh")) {
                const int k = as_int(*mi, "mesh");
                if (k < 0 || (size_t)k >= meshes.size())
                    bad("node mesh out of range");
                node_prims.back() = std::move(meshes[k]);   // a second user of this mesh gets none
                meshes[k].clear();
            }
            kids.emplace_back();
            if (const sj::Value *ch = n.get("children"))
                for (size_t k = 0; k < ch->size(); k++)
                    kids.back().push_back(as_int(at(*ch, k, "children"), "child"));
            A.nodes.push_back(d);
        }
        // roots: the nodes nobody lists as a child, in order
        std::vector<char> root(A.nodes.size(), 1);
        for (const auto &ks : kids)
            for (int c : ks) {
                if (c < 0 || (size_t)c >= A.nodes.size())
                    bad("child node out of range");
                root[c] = 0;
            }
        for (size_t i = 0; i < A.nodes.size(); i++)
            if (root[i])
                A.roots.push_back((int32_t)i);
        // flatten primitives and children in node order
        for (size_t i = 0; i < A.nodes.size(); i++) {
            swegl_node_desc &d = A.nodes[i];
            d.first_primitive = (uint32_t)A.prims.size();
            d.n_primitives = (uint32_t)node_prims[i].size();
            for (Prim &P : node_prims[i]) {
                const uint32_t nv = (uint32_t)(P.pos.size() / 3);
                for (uint32_t v : P.idx)
                    if (v >= nv)
                        bad("index " + std::to_string(v) + " past the primitive's vertices");
                A.pos.push_back(std::move(P.pos));
                A.nor.push_back(std::move(P.nor));
                A.uv.push_back(std::move(P.uv));
                A.idx.push_back(std::move(P.idx));
                swegl_primitive_desc pd;
                memset(&pd, 0, sizeof pd);
                pd.n_vertices = nv;
                pd.n_indices = (uint32_t)A.idx.back().size();
                pd.mode = P.mode;
                pd.material_id = P.material;
                A.prims.push_back(pd);
            }
            d.first_child = (uint32_t)A.children.size();
            d.n_children = (uint32_t)kids[i].size();
            A.children.insert(A.children.end(), kids[i].begin(), kids[i].end());
        }
    }

    // gltf.cpp:320-418
    void load_animations()
    {
        const sj::Value *ja = j.get("animations");
        if (!ja)
            return;
        for (const sj::Value &an : ja->arr) {
            swegl_animation anim{0.0f, 0, nullptr};
            const sj::Value &samplers = need(an, "samplers", "animation");
            std::vector<swegl_anim_channel> chans;
            for (const sj::Value &ch : need(an, "channels", "animation").arr) {
                const sj::Value &target = need(ch, "target", "channel");
                swegl_anim_channel c{0, SWEGL_PATH_NONE, 0, nullptr};
                c.node = as_int(need(target, "node", "channel target"), "target node");
                if (c.node < 0 || (size_t)c.node >= A.nodes.size())
                    bad("animation target node out of range");
                const sj::Value &path = need(target, "path", "channel target");
                const std::string p = path.is_str() ? path.str : std::string();
                c.path = p == "scale" ? SWEGL_PATH_SCALE : p == "rotation" ? SWEGL_PATH_ROTATION
                       : p == "translation" ? SWEGL_PATH_TRANSLATION : SWEGL_PATH_NONE;
                const int si = as_int(need(ch, "sampler", "channel"), "sampler");
                const sj::Value &s = at(samplers, si, "sampler");
                const Accessor &times = accessor(need(s, "input", "sampler"), "sampler input");
                const Accessor &vals = accessor(need(s, "output", "sampler"), "sampler output");
                std::vector<swegl_anim_step> st;
                for (int i = 0; i < times.count; i++) {
                    swegl_anim_step step{f32_at(times, (size_t)i * times.stride), {0, 0, 0, 1}};
                    const int nk = c.path == SWEGL_PATH_ROTATION ? 4 : c.path == SWEGL_PATH_NONE ? 0 : 3;
                    for (int k = 0; k < nk; k++)
                        step.value[k] = f32_at(vals, (size_t)i * vals.stride + 4 * k);
                    st.push_back(step);
                }
                if (st.empty())
                    bad("animation channel without keyframes");
                std::stable_sort(st.begin(), st.end(), [](const swegl_anim_step &a, const swegl_anim_step &b) {
                    return a.time < b.time;
                });
                anim.end_time = std::max(anim.end_time, st.back().time);
                A.steps.push_back(std::move(st));
                c.n_steps = (uint32_t)A.steps.back().size();
                chans.push_back(c);
            }
            A.channels.push_back(std::move(chans));
            A.animations.push_back(anim);
        }
    }

    void finish()
    {
        for (size_t i = 0; i < A.prims.size(); i++) {
            A.prims[i].positions = A.pos[i].data();
            A.prims[i].normals = A.nor[i].data();
            A.prims[i].tex_coords = A.uv[i].data();
            A.prims[i].indices = A.idx[i].data();
        }
        for (size_t i = 0; i < A.textures.size(); i++)
            A.textures[i].bitmap = A.bitmaps[i].data();
        size_t s = 0;
        for (size_t a = 0; a < A.animations.size(); a++) {
            for (swegl_anim_channel &c : A.channels[a])
                c.steps = A.steps[s++].data();
            A.animations[a].n_channels = (uint32_t)A.channels[a].size();
            A.animations[a].channels = A.channels[a].data();
        }
    }

    void parse_json(const uint8_t *b, const uint8_t *e)
    {
        std::string err;
        if (!sj::parse((const char *)b, (const char *)e, j, err))
            bad(filename + ": " + err);
        if (!j.is_obj())
            bad(filename + ": the document is not a JSON object");
    }

    // gltf.cpp:420-464
    void load()
    {
        root = parent_dir(filename);
        const std::string l = lower(filename);
        Bytes bin;
        bool have_bin = false;
        if (ends_with(l, "glb")) {
            files.push_back(read_file(filename));
            const std::vector<uint8_t> &g = files.back();
            if (g.size() < 12)
                bad("GLB shorter than its header");
            uint32_t length;
            memcpy(&length, g.data() + 8, 4);
            if (length > g.size())
                bad("GLB header length past the end of the file");
            std::vector<Bytes> chunks;
            for (size_t i = 12; i < length;) {
                if (i + 8 > length)
                    bad("GLB chunk header past the end of the file");
                uint32_t clen;
                memcpy(&clen, g.data() + i, 4);
                if (clen > length - i - 8)
                    bad("GLB chunk past the end of the file");
                chunks.push_back(Bytes{g.data() + i + 8, clen, clen});
                i += 8 + (size_t)clen;
            }
            if (chunks.empty())
                bad("GLB without a JSON chunk");
            parse_json(chunks[0].p, chunks[0].p + chunks[0].len);
            if (chunks.size() > 1) {
                bin = chunks[1];
                have_bin = true;
            }
        } else if (ends_with(l, "gltf")) {
            files.push_back(read_file(filename));
            parse_json(files.back().data(), files.back().data() + files.back().size());
        } else {
            bad("not a .glb or .gltf file: " + filename, SWEGL_ENOTSUP);
        }
        load_buffers(have_bin ? &bin : nullptr);
        load_views();
        load_accessors();
        load_images();
        load_materials();
        std::vector<std::vector<Prim>> meshes = load_meshes();
        load_nodes(meshes);
        load_animations();
        finish();
    }
};

// animation_channel_t::get_steps (model.hpp:109-121): the keyframes around t
void steps_around(const swegl_anim_channel &c, float t, const swegl_anim_step *&before,
                  const swegl_anim_step *&after)
{
    const swegl_anim_step *b = c.steps, *e = c.steps + c.n_steps;
    const swegl_anim_step *it =
        std::lower_bound(b, e, t, [](const swegl_anim_step &s, float x) { return s.time < x; });
    if (it == e)
        before = after = e - 1;
    else if (it == b)
        before = after = b;
    else {
        before = it - 1;
        after = it;
    }
}

}  // namespace

extern "C" int swegl_asset_load(const char *filename, swegl_asset **out)
{
    if (!filename || !out)
        return sw_fail_msg(SWEGL_EINVAL, "null argument");
    *out = nullptr;
    std::unique_ptr<swegl_asset> a(new swegl_asset);
    try {
        Loader L(*a);
        L.filename = filename;
        L.load();
    } catch (const Error &e) {
        return sw_fail_msg(e.rc, e.msg);
    } catch (const std::bad_alloc &) {
        return sw_fail_msg(SWEGL_ENOMEM, "out of host memory reading the asset");
    }
    *out = a.release();
    return SWEGL_OK;
}

extern "C" int swegl_asset_scene(const swegl_asset *a, swegl_scene_desc *d)
{
    if (!a || !d)
        return sw_fail_msg(SWEGL_EINVAL, "null argument");
    memset(d, 0, sizeof *d);
    d->n_nodes = (uint32_t)a->nodes.size();
    d->nodes = a->nodes.data();
    d->n_root_nodes = (uint32_t)a->roots.size();
    d->root_nodes = a->roots.data();
    d->n_child_indices = (uint32_t)a->children.size();
    d->child_indices = a->children.data();
    d->n_primitives = (uint32_t)a->prims.size();
    d->primitives = a->prims.data();
    d->n_materials = (uint32_t)a->materials.size();
    d->materials = a->materials.data();
    d->default_material = swegl_material_desc{0xFFFFFFFFu, 1.0f, 1.0f, -1, 0};
    d->n_textures = (uint32_t)a->textures.size();
    d->textures = a->textures.data();
    d->sun_direction[1] = -1.0f;
    return SWEGL_OK;
}

extern "C" int swegl_asset_animations(const swegl_asset *a, const swegl_animation **anims, uint32_t *n)
{
    if (!a || !anims || !n)
        return sw_fail_msg(SWEGL_EINVAL, "null argument");
    *anims = a->animations.data();
    *n = (uint32_t)a->animations.size();
    return SWEGL_OK;
}

extern "C" void swegl_asset_destroy(swegl_asset *a) { delete a; }

// scene_t::animate (model.hpp:146-177)
extern "C" int swegl_animate(const swegl_animation *anims, uint32_t n_anims, float elapsed_seconds,
                             swegl_node_desc *nodes, uint32_t n_nodes)
{
    if ((n_anims && !anims) || (n_nodes && !nodes))
        return sw_fail_msg(SWEGL_EINVAL, "null argument");
    for (uint32_t a = 0; a < n_anims; a++)
        for (uint32_t k = 0; k < anims[a].n_channels; k++) {
            const swegl_anim_channel &c = anims[a].channels[k];
            if (c.node < 0 || (uint32_t)c.node >= n_nodes || !c.n_steps || !c.steps)
                return sw_fail_msg(SWEGL_EINVAL, "animation channel: node out of range or no keyframes");
        }
    for (uint32_t a = 0; a < n_anims; a++) {
        const float rel = std::fmod(elapsed_seconds, anims[a].end_time);
        for (uint32_t k = 0; k < anims[a].n_channels; k++) {
            const swegl_anim_channel &c = anims[a].channels[k];
            const swegl_anim_step *b, *e;
            steps_around(c, rel, b, e);
            float f[4];
            if (e->time == b->time) {
                memcpy(f, b->value, sizeof f);
            } else {
                const float wb = (e->time - rel) / (e->time - b->time);
                const float we = (rel - b->time) / (e->time - b->time);
                for (int i = 0; i < 4; i++)
                    f[i] = b->value[i] * wb + e->value[i] * we;
            }
            swegl_node_desc &nd = nodes[c.node];
            if (c.path == SWEGL_PATH_ROTATION) {
                // vec4f_t::normalize (vec2f.hpp:61-71)
                const float len = (float)std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3]);
                if (len != 0)
                    for (int i = 0; i < 4; i++)
                        f[i] = f[i] / len;
                from_quaternion(f[0], f[1], f[2], f[3], nd.rotation);
            } else if (c.path == SWEGL_PATH_TRANSLATION) {
                memcpy(nd.translation, f, 3 * sizeof(float));
            } else if (c.path == SWEGL_PATH_SCALE) {
                memcpy(nd.scale, f, 3 * sizeof(float));
            }
        }
    }
    return SWEGL_OK;
}

extern "C" void swegl_matrix_from_quaternion(float m[16], float q0, float q1, float q2, float q3)
{
    from_quaternion(q0, q1, q2, q3, m);
}

extern "C" int swegl_image_read(const char *filename, uint32_t **pixels, uint32_t *w, uint32_t *h)
{
    if (!filename || !pixels || !w || !h)
        return sw_fail_msg(SWEGL_EINVAL, "null argument");
    *pixels = nullptr;
    try {
        const std::string l = lower(filename);
        const bool png = ends_with(l, "png");
        if (!png && !ends_with(l, "bmp")) {
            if (ends_with(l, "jpg") || ends_with(l, "jpeg"))
                bad("JPEG textures need a JPEG decoder, which this build does not have", SWEGL_ENOTSUP);
            bad(std::string("image format of ") + filename + " is not read by the reference", SWEGL_ENOTSUP);
        }
        std::vector<uint8_t> bytes = read_file(filename);
        std::vector<uint32_t> px;
        int iw = 0, ih = 0;
        std::string err;
        const bool ok = png ? sw_img::decode_png(bytes.data(), bytes.size(), px, iw, ih, err)
                            : sw_img::decode_bmp(bytes.data(), bytes.size(), px, iw, ih, err);
        if (!ok)
            bad(std::string(filename) + ": " + err);
        uint32_t *buf = (uint32_t *)malloc(px.size() * sizeof(uint32_t));
        if (!buf)
            bad("out of host memory", SWEGL_ENOMEM);
        memcpy(buf, px.data(), px.size() * sizeof(uint32_t));
        *pixels = buf;
        *w = (uint32_t)iw;
        *h = (uint32_t)ih;
    } catch (const Error &e) {
        return sw_fail_msg(e.rc, e.msg);
    }
    return SWEGL_OK;
}

extern "C" void swegl_image_free(uint32_t *pixels) { free(pixels); }

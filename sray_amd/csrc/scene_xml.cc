/*
 * scene_xml.cc — reference-format scene XML: a small self-contained parser
 * (the reference uses expat, src/xmltree.c:180-233) and the loader that turns
 * the document into a flat scene description, following Scene::load
 * (src/scene.cc:93-183) and the per-type load_xml methods it calls.
 * Also a writer that emits the same format (used by the generators so that
 * every synthetic workload exists as a scene file the reference could read).
 */
#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>

#include <unistd.h>

#include "host.h"

namespace sray {

/* ================================================================ */
/* XML DOM                                                            */
/* ================================================================ */
const XmlAttr *XmlNode::get(const char *n) const
{
    for (const XmlAttr &a : attr)
        if (a.name == n) return &a;
    return nullptr;
}

/* Attribute typing with the outcome of xml_create_attr, src/xmltree.c:51-99
 * (Q16).  The value is split into whitespace-separated words.  It counts as
 * numeric only when its first character is a digit, or a sign followed by a
 * digit, so ".5" and " 1" stay strings.  A numeric value is a VEC when it has
 * two or more words (the first four become components, the rest default to
 * 1.0), otherwise FLT when a '.' appears anywhere in it and INT when none
 * does.  Each component is the leading number of its word, as strtod reads
 * it ("1,2" gives 1). */
namespace {
bool opens_number(const std::string &v)
{
    auto digit = [&](size_t k) { return k < v.size() && isdigit((unsigned char)v[k]); };
    if (digit(0)) return true;
    return !v.empty() && (v[0] == '+' || v[0] == '-') && digit(1);
}

std::vector<std::string> words_of(const std::string &v, size_t max_words)
{
    std::vector<std::string> w;
    size_t k = 0;
    while (w.size() < max_words) {
        while (k < v.size() && isspace((unsigned char)v[k])) k++;
        if (k == v.size()) break;
        size_t e = k;
        while (e < v.size() && !isspace((unsigned char)v[e])) e++;
        w.push_back(v.substr(k, e - k));
        k = e;
    }
    return w;
}
} // namespace

XmlAttr xml_make_attr(const std::string &name, const std::string &val)
{
    XmlAttr a;
    a.name = name;
    a.str = val;
    if (!opens_number(val)) return a;
    const std::vector<std::string> w = words_of(val, 4);
    a.ival = (int)strtol(w[0].c_str(), nullptr, 10);
    a.fval = (float)strtod(w[0].c_str(), nullptr);
    for (float &c : a.vval) c = 1.0f;
    for (size_t k = 0; k < w.size(); k++) a.vval[k] = (float)strtod(w[k].c_str(), nullptr);
    if (w.size() >= 2)
        a.type = ATYPE_VEC;
    else
        a.type = val.find('.') != std::string::npos ? ATYPE_FLT : ATYPE_INT;
    return a;
}

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;
    explicit Parser(const std::string &t) : s(t) {}

    bool eof() const { return i >= s.size(); }
    void skip_ws()
    {
        while (!eof() && isspace((unsigned char)s[i])) i++;
    }
    bool starts(const char *p) const { return s.compare(i, strlen(p), p) == 0; }
    bool skip_until(const char *p)
    {
        size_t k = s.find(p, i);
        if (k == std::string::npos) return false;
        i = k + strlen(p);
        return true;
    }
    static std::string unescape(const std::string &v)
    {
        std::string o;
        for (size_t k = 0; k < v.size(); k++) {
            if (v[k] != '&') {
                o += v[k];
                continue;
            }
            size_t e = v.find(';', k);
            if (e == std::string::npos) {
                o += v[k];
                continue;
            }
            std::string ent = v.substr(k + 1, e - k - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#') o += (char)strtol(ent.c_str() + 1 + (ent[1] == 'x'), nullptr, ent[1] == 'x' ? 16 : 10);
            else o += "&" + ent + ";";
            k = e;
        }
        return o;
    }
    bool name(std::string &n)
    {
        size_t b = i;
        while (!eof() && (isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == ':' || s[i] == '.')) i++;
        n = s.substr(b, i - b);
        return !n.empty();
    }
    /* skip prolog, comments, processing instructions, doctype */
    void skip_misc()
    {
        for (;;) {
            skip_ws();
            if (starts("<?")) {
                if (!skip_until("?>")) return;
            } else if (starts("<!--")) {
                if (!skip_until("-->")) return;
            } else if (starts("<!")) {
                if (!skip_until(">")) return;
            } else return;
        }
    }
    bool element(XmlNode &n, int depth)
    {
        if (depth > 256) {
            err = "nesting too deep";
            return false;
        }
        if (eof() || s[i] != '<') {
            err = "expected '<'";
            return false;
        }
        i++;
        if (!name(n.name)) {
            err = "bad element name";
            return false;
        }
        for (;;) {
            skip_ws();
            if (eof()) {
                err = "unterminated tag <" + n.name + ">";
                return false;
            }
            if (starts("/>")) {
                i += 2;
                return true;
            }
            if (s[i] == '>') {
                i++;
                break;
            }
            std::string an;
            if (!name(an)) {
                err = "bad attribute in <" + n.name + ">";
                return false;
            }
            skip_ws();
            if (eof() || s[i] != '=') {
                err = "expected '=' after " + an;
                return false;
            }
            i++;
            skip_ws();
            if (eof() || (s[i] != '"' && s[i] != '\'')) {
                err = "expected quoted value for " + an;
                return false;
            }
            char q = s[i++];
            size_t e = s.find(q, i);
            if (e == std::string::npos) {
                err = "unterminated attribute value";
                return false;
            }
            n.attr.push_back(xml_make_attr(an, unescape(s.substr(i, e - i))));
            i = e + 1;
        }
        /* content */
        for (;;) {
            size_t lt = s.find('<', i);
            if (lt == std::string::npos) {
                err = "unterminated element <" + n.name + ">";
                return false;
            }
            i = lt;
            if (starts("</")) {
                i += 2;
                std::string en;
                name(en);
                skip_ws();
                if (eof() || s[i] != '>' || en != n.name) {
                    err = "mismatched </" + en + "> for <" + n.name + ">";
                    return false;
                }
                i++;
                return true;
            }
            if (starts("<!--")) {
                if (!skip_until("-->")) {
                    err = "unterminated comment";
                    return false;
                }
                continue;
            }
            if (starts("<![CDATA[")) {
                if (!skip_until("]]>")) {
                    err = "unterminated CDATA";
                    return false;
                }
                continue;
            }
            if (starts("<?")) {
                if (!skip_until("?>")) {
                    err = "unterminated PI";
                    return false;
                }
                continue;
            }
            n.chld.emplace_back();
            if (!element(n.chld.back(), depth + 1)) return false;
        }
    }
};

} // namespace

bool xml_parse(const std::string &text, XmlNode &root, std::string &err)
{
    Parser p(text);
    p.skip_misc();
    if (p.eof()) {
        err = "empty document";
        return false;
    }
    if (!p.element(root, 0)) {
        err = p.err;
        return false;
    }
    return true;
}

/* ================================================================ */
/* transforms: XFormNode::load_xml (src/anim.cc:68-121)               */
/* ================================================================ */
namespace {

void load_xform(const XmlNode &n, XFormTracks &xf)
{
    int msec = 0;
    const XmlAttr *a;
    if ((a = n.get("interpolate"))) {
        if (a->str == "linear") xf.interp = SR_INTERP_LINEAR;
        else if (a->str == "step") xf.interp = SR_INTERP_STEP;
        else if (a->str == "cubic") xf.interp = SR_INTERP_CUBIC;
    }
    if ((a = n.get("extrapolate"))) {
        if (a->str == "clamp") xf.extrap = SR_EXTRAP_CLAMP;
        else if (a->str == "repeat") xf.extrap = SR_EXTRAP_REPEAT;
        else if (a->str == "pingpong") xf.extrap = SR_EXTRAP_PINGPONG;
    }
    if ((a = n.get("time")) && a->type == ATYPE_INT) msec = a->ival;
    if ((a = n.get("pos")) && a->type == ATYPE_VEC) xf.set_pos(msec, a->vval[0], a->vval[1], a->vval[2]);
    if ((a = n.get("rot")) && a->type == ATYPE_VEC) {
        /* Quaternion(s, x, y, z), normalised */
        double q[4] = {a->vval[0], a->vval[1], a->vval[2], a->vval[3]};
        double l = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        if (l > 0)
            for (double &v : q) v /= l;
        xf.set_rot(msec, q, a->vval, 0);
    }
    if ((a = n.get("axisrot")) && a->type == ATYPE_VEC) {
        double q[4];
        quat_from_axis_angle(a->vval[1], a->vval[2], a->vval[3], a->vval[0] * M_PI / 180.0, q);
        xf.set_rot(msec, q, a->vval, 1);
    }
    if ((a = n.get("scale")) && a->type == ATYPE_VEC) xf.set_scale(msec, a->vval[0], a->vval[1], a->vval[2]);
    if ((a = n.get("pivot")) && a->type == ATYPE_VEC)
        for (int k = 0; k < 3; k++) xf.pivot[k] = a->vval[k];
}

void load_xforms(const XmlNode &n, XFormTracks &xf)
{
    for (const XmlNode &c : n.chld)
        if (c.name == "xform") load_xform(c, xf);
}

/* ================================================================ */
/* scene loader                                                       */
/* ================================================================ */
sray_material_desc default_material()
{
    /* Material::Material, src/material.cc:79-89 (shade_undef shader) */
    sray_material_desc m;
    std::memset(&m, 0, sizeof m);
    m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = m.diffuse[3] = 1.0f;
    m.specular[3] = 1.0f;
    m.shininess = 1.0f; /* missing attribute -> def_attr (1,1,1,1) */
    m.reflect = 0.0f;
    m.refract = 0.0f;
    m.ior = 1.5f;
    m.flags = SRAY_MAT_HAS_SPECULAR | SRAY_MAT_HAS_REFLECT | SRAY_MAT_HAS_REFRACT;
    return m;
}

/* find_file, src/datapath.c:56-77: the first "<dir>/<name>" that exists
 * over the search list — the scene file's directory (src/sray.cc:103-112;
 * "." for a scene given as text, src/sray.cc:116-118, or a scene file named
 * without a directory), then the entries of SRAY_DATA_PATH, ':'-separated.
 * An absolute name is taken as it is. */
std::string find_data_file(const std::string &name, const std::string &scene_dir)
{
    if (!name.empty() && name[0] == '/') return access(name.c_str(), F_OK) == 0 ? name : std::string();
    std::vector<std::string> dirs;
    dirs.push_back(scene_dir.empty() ? std::string(".") : scene_dir);
    if (const char *env = getenv("SRAY_DATA_PATH")) {
        std::string e(env);
        size_t at = 0;
        while (true) {
            size_t c = e.find(':', at);
            dirs.push_back(e.substr(at, c == std::string::npos ? std::string::npos : c - at));
            if (c == std::string::npos) break;
            at = c + 1;
        }
    }
    for (const std::string &d : dirs) {
        std::string p = d + "/" + name;
        if (access(p.c_str(), F_OK) == 0) return p;
    }
    return std::string();
}

/* Texture::load (src/texture.cc:28-37) for a "map" attribute: 1 + the
 * scene's index of the texture (one per file, shared by the attributes
 * that name it; linear filter, repeat wrap, identity matrix, Texture(),
 * src/texture.cc:21-25), or 0 when it cannot be loaded — the reference
 * then reports it and the attribute keeps its plain colour
 * (src/material.cc:64-71) */
int load_texture(const std::string &name, const std::string &scene_dir, HostScene &hs)
{
    const std::string path = find_data_file(name, scene_dir);
    if (path.empty()) {
        fprintf(stderr, "failed to load texture: %s\n", name.c_str());
        return 0;
    }
    for (size_t i = 0; i < hs.texture_paths.size(); i++)
        if (hs.texture_paths[i] == path) return (int)i + 1;
    int w, h;
    std::vector<float> px;
    std::string ierr;
    if (!read_image_rgbaf(path, w, h, px, ierr)) {
        fprintf(stderr, "failed to load texture: %s (%s)\n", name.c_str(), ierr.c_str());
        return 0;
    }
    sray_texture_desc t;
    std::memset(&t, 0, sizeof t);
    t.width = w;
    t.height = h;
    t.filter = SRAY_TEX_FILTER_LINEAR;
    t.wrap = SRAY_TEX_WRAP_REPEAT;
    t.xform[0] = t.xform[5] = t.xform[10] = 1.0f;
    hs.textures.push_back(t);
    hs.texels.push_back(std::move(px));
    hs.texture_paths.push_back(path);
    hs.texture_names.push_back(name);
    return (int)hs.textures.size();
}

/* Material::load_xml + MatAttrib::load_xml, src/material.cc:36-124 */
bool load_material(const XmlNode &n, HostScene &hs, const std::string &scene_dir, std::string &err)
{
    const XmlAttr *a = n.get("name");
    if (!a) {
        err = "invalid material: name attribute not found";
        return false;
    }
    sray_material_desc m = default_material();
    std::string name = a->str;
    if ((a = n.get("shader")) && a->str == "phong") m.flags |= SRAY_MAT_PHONG;
    for (const XmlNode &c : n.chld) {
        if (c.name != "mattr") continue;
        const XmlAttr *an = c.get("name"), *av = c.get("value");
        if (!an || !av) continue; /* MatAttrib::load_xml fails -> attribute skipped */
        float col[4];
        if (av->type == ATYPE_VEC) {
            col[0] = av->vval[0];
            col[1] = av->vval[1];
            col[2] = av->vval[2];
            col[3] = 1.0f;
        } else if (av->type == ATYPE_FLT || av->type == ATYPE_INT) {
            col[0] = col[1] = col[2] = av->fval;
            col[3] = 1.0f;
        } else continue;
        const std::string &nm = an->str;
        int slot = -1;
        if (nm == "diffuse") {
            std::memcpy(m.diffuse, col, sizeof col);
            slot = SRAY_MAP_DIFFUSE;
        } else if (nm == "specular") {
            std::memcpy(m.specular, col, sizeof col);
            m.flags |= SRAY_MAT_HAS_SPECULAR;
            slot = SRAY_MAP_SPECULAR;
        } else if (nm == "shininess") {
            m.shininess = col[0];
            slot = SRAY_MAP_SHININESS;
        } else if (nm == "reflect") {
            m.reflect = col[0];
            m.flags |= SRAY_MAT_HAS_REFLECT;
            slot = SRAY_MAP_REFLECT;
        } else if (nm == "refract") {
            m.refract = col[0];
            m.flags |= SRAY_MAT_HAS_REFRACT;
            slot = SRAY_MAP_REFRACT;
        } else if (nm == "ior") {
            m.ior = col[0];
            slot = SRAY_MAP_IOR;
        } else if (nm == "normal") {
            m.normal = col[0];
            m.flags |= SRAY_MAT_HAS_NORMAL;
            slot = SRAY_MAP_NORMAL;
        }
        /* set_attribute replaces colour and texture together (src/material.cc:150-161) */
        if (slot >= 0) {
            const XmlAttr *am = c.get("map");
            m.map[slot] = am ? load_texture(am->str, scene_dir, hs) : 0;
        }
    }
    hs.materials.push_back(m);
    hs.material_names.push_back(name);
    return true;
}

/* Mesh::load_xml, src/meshload.cc:55-157 */
bool load_mesh(const XmlNode &obj, HostScene &hs, sray_object_desc &o, std::string &err)
{
    const XmlNode *mn = nullptr;
    for (const XmlNode &c : obj.chld)
        if (c.name == "mesh") {
            mn = &c;
            break;
        }
    if (!mn) {
        err = "<mesh> not found in mesh object";
        return false;
    }
    static const char *el[4] = {"vertex", "normal", "texcoord", "tangent"};
    struct VE {
        float v[3];
        int id;
    };
    std::vector<VE> ve[4];
    bool ordered[4] = {true, true, true, true};
    int cur[4] = {0, 0, 0, 0};
    struct FR {
        int vi[3], ni[3], gi[3], ti[3];
    };
    std::vector<FR> faces;
    for (const XmlNode &c : mn->chld) {
        int e = -1;
        for (int k = 0; k < 4; k++)
            if (c.name == el[k]) e = k;
        if (e >= 0) {
            VE x;
            const XmlAttr *a;
            x.id = ((a = c.get("id")) && a->type == ATYPE_INT) ? a->ival : cur[e];
            x.v[0] = x.v[1] = x.v[2] = 0;
            if ((a = c.get("val")) && a->type == ATYPE_VEC) {
                x.v[0] = a->vval[0];
                x.v[1] = a->vval[1];
                x.v[2] = a->vval[2];
            }
            if (x.id != cur[e]++) ordered[e] = false;
            ve[e].push_back(x);
        } else if (c.name == "face") {
            FR f;
            std::memset(&f, 0, sizeof f);
            int i = 0;
            for (const XmlNode &r : c.chld) {
                if (i >= 3) break;
                if (r.name != "vref") continue;
                const XmlAttr *a;
                f.vi[i] = ((a = r.get("vertex")) && a->type == ATYPE_INT) ? a->ival : 0;
                f.ni[i] = ((a = r.get("normal")) && a->type == ATYPE_INT) ? a->ival : 0;
                f.gi[i] = ((a = r.get("tangent")) && a->type == ATYPE_INT) ? a->ival : 0;
                f.ti[i] = ((a = r.get("texcoord")) && a->type == ATYPE_INT) ? a->ival : 0;
                i++;
            }
            faces.push_back(f);
        }
    }
    for (int k = 0; k < 4; k++) {
        if (!ordered[k])
            std::stable_sort(ve[k].begin(), ve[k].end(), [](const VE &a, const VE &b) { return a.id < b.id; });
        if (!ve[k].empty() && ve[k].back().id != (int)ve[k].size() - 1) {
            err = "error loading mesh: missing vertex elements";
            return false;
        }
    }
    bool has_n = !ve[1].empty(), has_t = !ve[2].empty(), has_g = !ve[3].empty();
    /* attribute arrays are scene-wide: start them when the first mesh with
     * that attribute appears, zero-filled for earlier triangles */
    size_t ntri0 = hs.tri_pos.size() / 9;
    auto ensure = [&](bool &flag, std::vector<float> &arr, size_t per) {
        if (!flag) {
            flag = true;
            arr.assign(ntri0 * per, 0.0f);
        }
    };
    if (has_n) ensure(hs.has_normal, hs.tri_normal, 9);
    if (has_g) ensure(hs.has_tangent, hs.tri_tangent, 9);
    if (has_t) ensure(hs.has_texcoord, hs.tri_texcoord, 6);
    sray_mesh_desc md;
    md.tri_start = (uint32_t)ntri0;
    md.tri_count = (uint32_t)faces.size();
    for (const FR &f : faces) {
        for (int j = 0; j < 3; j++) {
            if (f.vi[j] < 0 || f.vi[j] >= (int)ve[0].size()) {
                err = "mesh face references a missing vertex";
                return false;
            }
            for (int k = 0; k < 3; k++) hs.tri_pos.push_back(ve[0][f.vi[j]].v[k]);
        }
        if (hs.has_normal)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                    hs.tri_normal.push_back(has_n && f.ni[j] >= 0 && f.ni[j] < (int)ve[1].size() ? ve[1][f.ni[j]].v[k] : 0.0f);
        if (hs.has_tangent)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                    hs.tri_tangent.push_back(has_g && f.gi[j] >= 0 && f.gi[j] < (int)ve[3].size() ? ve[3][f.gi[j]].v[k] : 0.0f);
        if (hs.has_texcoord)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 2; k++)
                    hs.tri_texcoord.push_back(has_t && f.ti[j] >= 0 && f.ti[j] < (int)ve[2].size() ? ve[2][f.ti[j]].v[k] : 0.0f);
    }
    o.mesh = (int32_t)hs.meshes.size();
    hs.meshes.push_back(md);
    return true;
}

/* load_object + Object::load_xml + Sphere/Cylinder/Mesh::load_xml,
 * src/scene.cc:797-816, src/object.cc:116-135, 34-77 */
bool load_object(const XmlNode &n, HostScene &hs, std::string &err)
{
    const XmlAttr *a = n.get("type");
    if (!a) {
        err = "invalid object: no type attribute";
        return false;
    }
    sray_object_desc o;
    std::memset(&o, 0, sizeof o);
    o.material = -1;
    o.mesh = -1;
    o.radius = 1.0f;
    o.start = 0.0f;
    o.end = 1.0f;
    if (a->str == "sphere") o.type = SRAY_OBJ_SPHERE;
    else if (a->str == "cylinder") o.type = SRAY_OBJ_CYLINDER;
    else if (a->str == "mesh") o.type = SRAY_OBJ_MESH;
    else {
        err = "invalid object: unknown type: " + a->str;
        return false;
    }
    XFormTracks xf;
    for (const XmlNode &c : n.chld) {
        if (c.name == "matref") {
            const XmlAttr *mn = c.get("name");
            if (!mn) {
                err = "invalid material reference: no name attribute";
                return false;
            }
            int found = -1;
            for (size_t k = 0; k < hs.material_names.size(); k++)
                if (hs.material_names[k] == mn->str) found = (int)k;
            if (found < 0) {
                err = "material " + mn->str + " referenced by object does not exist";
                return false;
            }
            o.material = found;
        } else if (c.name == "xform") {
            load_xform(c, xf);
        }
    }
    if (o.type == SRAY_OBJ_SPHERE || o.type == SRAY_OBJ_CYLINDER) {
        const char *tag = o.type == SRAY_OBJ_SPHERE ? "sphere" : "cylinder";
        bool res = false;
        for (const XmlNode &c : n.chld) {
            if (c.name != tag) continue;
            const XmlAttr *r;
            if ((r = c.get("rad"))) o.radius = r->fval;
            if (o.type == SRAY_OBJ_CYLINDER) {
                if ((r = c.get("start"))) o.start = r->fval;
                if ((r = c.get("end"))) o.end = r->fval;
            }
            res = true;
        }
        if (!res) {
            err = std::string("object has no <") + tag + "> element";
            return false;
        }
    } else if (!load_mesh(n, hs, o, err)) {
        return false;
    }
    /* the matrices come from the tracks (HostScene::set_frame) */
    hs.objects.push_back(o);
    hs.obj_tracks.push_back(xf);
    return true;
}

/* load_light + Light/SphLight/BoxLight::load_xml, src/light.cc:39-65,236-255,310-329 */
bool load_light(const XmlNode &n, HostScene &hs, std::string &err)
{
    const XmlAttr *a = n.get("type");
    if (!a) {
        err = "invalid light: no type attribute";
        return false;
    }
    sray_light_desc l;
    std::memset(&l, 0, sizeof l);
    if (a->str == "point") l.type = SRAY_LIGHT_POINT;
    else if (a->str == "sphere") l.type = SRAY_LIGHT_SPHERE;
    else if (a->str == "box") l.type = SRAY_LIGHT_BOX;
    else {
        err = "invalid light: unknown type: " + a->str;
        return false;
    }
    l.color[0] = l.color[1] = l.color[2] = l.color[3] = 1.0f;
    if ((a = n.get("color"))) {
        if (a->type == ATYPE_VEC) {
            l.color[0] = a->vval[0];
            l.color[1] = a->vval[1];
            l.color[2] = a->vval[2];
        } else if (a->type == ATYPE_INT || a->type == ATYPE_FLT) {
            l.color[0] = l.color[1] = l.color[2] = a->fval;
        }
    }
    l.radius = 1.0f;
    l.dim[0] = l.dim[1] = l.dim[2] = 1.0f;
    XFormTracks xf;
    load_xforms(n, xf);
    for (const XmlNode &c : n.chld) {
        if (l.type == SRAY_LIGHT_SPHERE && c.name == "sphere") {
            const XmlAttr *r = c.get("rad");
            if (!r) {
                err = "sphere light without rad";
                return false;
            }
            l.radius = r->fval;
        }
        if (l.type == SRAY_LIGHT_BOX && c.name == "aabox") {
            const XmlAttr *mn = c.get("min"), *mx = c.get("max");
            if (!mn || mn->type != ATYPE_VEC || !mx || mx->type != ATYPE_VEC) {
                err = "invalid <aabox> in box light";
                return false;
            }
            for (int k = 0; k < 3; k++) l.dim[k] = mx->vval[k] - mn->vval[k];
        }
    }
    /* position and box matrix come from the tracks (HostScene::set_frame) */
    hs.lights.push_back(l);
    hs.light_tracks.push_back(xf);
    return true;
}

} // namespace

/* TargetCamera::get_matrix, src/camera.cc:216-234 */
void target_camera_matrix(const double pos[3], const double tg[3], double roll, double m[12])
{
    double dx = tg[0] - pos[0], dy = tg[1] - pos[1], dz = tg[2] - pos[2];
    double dl = sqrt(dx * dx + dy * dy + dz * dz);
    dx /= dl;
    dy /= dl;
    dz /= dl;
    /* right = normalize(up x dir), up = dir x right */
    double ux = 0, uy = 1, uz = 0;
    double rx = uy * dz - uz * dy, ry = uz * dx - ux * dz, rz = ux * dy - uy * dx;
    double rl = sqrt(rx * rx + ry * ry + rz * rz);
    rx /= rl;
    ry /= rl;
    rz /= rl;
    ux = dy * rz - dz * ry;
    uy = dz * rx - dx * rz;
    uz = dx * ry - dy * rx;
    if (roll != 0.0) {
        double R[12], q[4];
        quat_from_axis_angle(dx, dy, dz, roll, q);
        quat_to_matrix(q, R);
        double nux = R[0] * ux + R[1] * uy + R[2] * uz, nuy = R[4] * ux + R[5] * uy + R[6] * uz,
               nuz = R[8] * ux + R[9] * uy + R[10] * uz;
        double nrx = R[0] * rx + R[1] * ry + R[2] * rz, nry = R[4] * rx + R[5] * ry + R[6] * rz,
               nrz = R[8] * rx + R[9] * ry + R[10] * rz;
        ux = nux, uy = nuy, uz = nuz, rx = nrx, ry = nry, rz = nrz;
    }
    double M[12] = {rx, ux, dx, pos[0], ry, uy, dy, pos[1], rz, uz, dz, pos[2]};
    std::memcpy(m, M, sizeof M);
}

namespace {

/* load_camera + Camera/TargetCamera::load_xml + get_matrix,
 * src/camera.cc:39-67, 161-192, 216-234 */
bool load_camera(const XmlNode &n, HostScene &hs, std::string &err)
{
    const XmlAttr *a = n.get("type");
    if (!a) {
        err = "invalid camera: no type attribute";
        return false;
    }
    bool target = false;
    if (a->str == "target") target = true;
    else if (a->str != "free") {
        err = "invalid camera: unknown type: " + a->str;
        return false;
    }
    sray_camera_desc c = default_camera();
    if ((a = n.get("shutter"))) c.shutter = (int)a->fval;
    if ((a = n.get("fov")) && (a->type == ATYPE_FLT || a->type == ATYPE_INT)) c.vfov = (double)a->fval * M_PI / 180.0;
    XFormTracks xf, tx;
    load_xforms(n, xf);
    if (target) {
        const XmlNode *tn = nullptr;
        for (const XmlNode &ch : n.chld)
            if (ch.name == "target") {
                tn = &ch;
                break;
            }
        if (!tn) {
            err = "failed to load target camera, no <target> child";
            return false;
        }
        load_xforms(*tn, tx);
    }
    /* the matrix comes from the tracks (HostScene::set_frame) */
    hs.camera = c;
    hs.cam_tracks = xf;
    hs.cam_target_tracks = tx;
    hs.cam_has_target = target;
    return true;
}

} // namespace

/* Scene::load(FILE*), src/scene.cc:93-183: materials first, then env,
 * objects, lights and the (first) camera in document order */
bool load_scene_xml(const std::string &text, HostScene &hs, std::string &err, const std::string &scene_dir)
{
    XmlNode root;
    if (!xml_parse(text, root, err)) return false;
    hs = HostScene();
    hs.camera = default_camera();
    for (const XmlNode &n : root.chld)
        if (n.name == "material" && !load_material(n, hs, scene_dir, err)) return false;
    bool have_cam = false;
    for (const XmlNode &n : root.chld) {
        if (n.name == "env") {
            const XmlAttr *a;
            if ((a = n.get("ambient"))) {
                if (a->type == ATYPE_VEC) {
                    hs.env_ambient[0] = a->vval[0];
                    hs.env_ambient[1] = a->vval[1];
                    hs.env_ambient[2] = a->vval[2];
                    hs.env_ambient[3] = 1.0f;
                } else if (a->type == ATYPE_INT || a->type == ATYPE_FLT) {
                    hs.env_ambient[0] = hs.env_ambient[1] = hs.env_ambient[2] = a->fval;
                    hs.env_ambient[3] = 1.0f;
                }
            }
            if ((a = n.get("bgcolor")) && a->type == ATYPE_VEC) {
                hs.env_color[0] = a->vval[0];
                hs.env_color[1] = a->vval[1];
                hs.env_color[2] = a->vval[2];
                hs.env_color[3] = 1.0f;
            }
            if ((a = n.get("ior")) && a->type == ATYPE_FLT) hs.env_ior = a->fval;
        } else if (n.name == "object") {
            if (!load_object(n, hs, err)) return false;
        } else if (n.name == "light") {
            if (!load_light(n, hs, err)) return false;
        } else if (n.name == "camera") {
            HostScene tmp;
            if (!load_camera(n, hs.camera.max_dist != 0 && have_cam ? tmp : hs, err)) return false;
            have_cam = true;
        }
    }
    /* render(0): every transform at msec 0, rows over the frame interval */
    hs.has_tracks = true;
    return hs.set_frame(0, err);
}

/* ================================================================ */
/* writer                                                             */
/* ================================================================ */
namespace {
std::string f9(double v)
{
    char b[64];
    snprintf(b, sizeof b, "%.9g", (double)(float)v);
    /* keep numbers typed as floats by the loader (a '.' present) and
     * starting with a digit or sign (Q16) */
    std::string s = b;
    if (s.find('.') == std::string::npos && s.find('e') == std::string::npos && s.find("inf") == std::string::npos &&
        s.find("nan") == std::string::npos)
        s += ".0";
    return s;
}
std::string vec3(const float *v) { return f9(v[0]) + " " + f9(v[1]) + " " + f9(v[2]); }
std::string vec3d(const double *v) { return f9(v[0]) + " " + f9(v[1]) + " " + f9(v[2]); }

/* the keys of an XFormNode's tracks as <xform> elements (the values as
 * they were read, so the scene reads back to the same tracks) */
void write_tracks(std::ostringstream &o, const XFormTracks &xf, const char *ind)
{
    static const char *ip[] = {"step", "linear", "cubic"}, *ep[] = {"clamp", "repeat", "pingpong"};
    const bool pivot = xf.pivot[0] != 0 || xf.pivot[1] != 0 || xf.pivot[2] != 0;
    if (xf.interp != SR_INTERP_LINEAR || xf.extrap != SR_EXTRAP_CLAMP || pivot) {
        o << ind << "<xform interpolate=\"" << ip[xf.interp] << "\" extrapolate=\"" << ep[xf.extrap] << "\"";
        if (pivot) o << " pivot=\"" << vec3d(xf.pivot) << "\"";
        o << "/>\n";
    }
    for (const TrackKey3 &k : xf.pos) o << ind << "<xform time=\"" << k.time << "\" pos=\"" << vec3d(k.v) << "\"/>\n";
    for (const TrackKeyQ &k : xf.rot)
        o << ind << "<xform time=\"" << k.time << "\" " << (k.axisrot ? "axisrot" : "rot") << "=\"" << f9(k.raw[0]) << " "
          << f9(k.raw[1]) << " " << f9(k.raw[2]) << " " << f9(k.raw[3]) << "\"/>\n";
    for (const TrackKey3 &k : xf.scale)
        o << ind << "<xform time=\"" << k.time << "\" scale=\"" << vec3d(k.v) << "\"/>\n";
}
} // namespace

std::string write_scene_xml(const HostScene &hs)
{
    std::ostringstream o;
    o << "<scene>\n";
    o << "  <env ambient=\"" << vec3(hs.env_ambient) << "\"";
    if (hs.env_color[3] != 0.0f) o << " bgcolor=\"" << vec3(hs.env_color) << "\"";
    o << " ior=\"" << f9(hs.env_ior) << "\"/>\n";
    for (size_t i = 0; i < hs.materials.size(); i++) {
        const sray_material_desc &m = hs.materials[i];
        std::string nm = i < hs.material_names.size() ? hs.material_names[i] : ("mat" + std::to_string(i));
        o << "  <material name=\"" << nm << "\"" << ((m.flags & SRAY_MAT_PHONG) ? " shader=\"phong\"" : "") << ">\n";
        /* a mapped attribute names its texture as the scene it came from did */
        auto mp = [&](int k) -> std::string {
            const int t = m.map[k];
            if (t <= 0 || t > (int)hs.texture_names.size()) return "";
            return " map=\"" + hs.texture_names[t - 1] + "\"";
        };
        o << "    <mattr name=\"diffuse\" value=\"" << vec3(m.diffuse) << "\"" << mp(SRAY_MAP_DIFFUSE) << "/>\n";
        o << "    <mattr name=\"specular\" value=\"" << vec3(m.specular) << "\"" << mp(SRAY_MAP_SPECULAR) << "/>\n";
        o << "    <mattr name=\"shininess\" value=\"" << f9(m.shininess) << "\"" << mp(SRAY_MAP_SHININESS) << "/>\n";
        o << "    <mattr name=\"reflect\" value=\"" << f9(m.reflect) << "\"" << mp(SRAY_MAP_REFLECT) << "/>\n";
        o << "    <mattr name=\"refract\" value=\"" << f9(m.refract) << "\"" << mp(SRAY_MAP_REFRACT) << "/>\n";
        o << "    <mattr name=\"ior\" value=\"" << f9(m.ior) << "\"" << mp(SRAY_MAP_IOR) << "/>\n";
        if (m.flags & SRAY_MAT_HAS_NORMAL)
            o << "    <mattr name=\"normal\" value=\"" << f9(m.normal) << "\"" << mp(SRAY_MAP_NORMAL) << "/>\n";
        o << "  </material>\n";
    }
    for (size_t oi = 0; oi < hs.objects.size(); oi++) {
        const sray_object_desc &ob = hs.objects[oi];
        const char *ty = ob.type == SRAY_OBJ_SPHERE ? "sphere" : ob.type == SRAY_OBJ_CYLINDER ? "cylinder" : "mesh";
        o << "  <object type=\"" << ty << "\">\n";
        if (ob.material >= 0 && ob.material < (int)hs.material_names.size())
            o << "    <matref name=\"" << hs.material_names[ob.material] << "\"/>\n";
        /* generators only place objects by translation (+ optional motion) */
        float p0[3] = {ob.xform[3], ob.xform[7], ob.xform[11]};
        bool moving = (ob.motion[0] != 0 || ob.motion[1] != 0 || ob.motion[2] != 0) && ob.motion_t1 > ob.motion_t0;
        if (hs.has_tracks && oi < hs.obj_tracks.size()) {
            write_tracks(o, hs.obj_tracks[oi], "    ");
        } else if (moving) {
            float p1[3] = {p0[0] + ob.motion[0], p0[1] + ob.motion[1], p0[2] + ob.motion[2]};
            o << "    <xform time=\"" << ob.motion_t0 << "\" pos=\"" << vec3(p0) << "\"/>\n";
            o << "    <xform time=\"" << ob.motion_t1 << "\" pos=\"" << vec3(p1) << "\"/>\n";
        } else {
            o << "    <xform pos=\"" << vec3(p0) << "\"/>\n";
        }
        if (ob.type == SRAY_OBJ_SPHERE) o << "    <sphere rad=\"" << f9(ob.radius) << "\"/>\n";
        else if (ob.type == SRAY_OBJ_CYLINDER)
            o << "    <cylinder rad=\"" << f9(ob.radius) << "\" start=\"" << f9(ob.start) << "\" end=\"" << f9(ob.end)
              << "\"/>\n";
        else {
            const sray_mesh_desc &md = hs.meshes[ob.mesh];
            o << "    <mesh>\n";
            /* one vertex per triangle corner keeps the writer simple and exact */
            for (uint32_t f = 0; f < md.tri_count; f++)
                for (int c = 0; c < 3; c++) {
                    const float *v = &hs.tri_pos[(size_t)(md.tri_start + f) * 9 + c * 3];
                    o << "      <vertex id=\"" << (f * 3 + c) << "\" val=\"" << vec3(v) << "\"/>\n";
                }
            if (hs.has_normal)
                for (uint32_t f = 0; f < md.tri_count; f++)
                    for (int c = 0; c < 3; c++) {
                        const float *v = &hs.tri_normal[(size_t)(md.tri_start + f) * 9 + c * 3];
                        o << "      <normal id=\"" << (f * 3 + c) << "\" val=\"" << vec3(v) << "\"/>\n";
                    }
            if (hs.has_texcoord)
                for (uint32_t f = 0; f < md.tri_count; f++)
                    for (int c = 0; c < 3; c++) {
                        const float *v = &hs.tri_texcoord[(size_t)(md.tri_start + f) * 6 + c * 2];
                        o << "      <texcoord id=\"" << (f * 3 + c) << "\" val=\"" << f9(v[0]) << " " << f9(v[1])
                          << "\"/>\n";
                    }
            for (uint32_t f = 0; f < md.tri_count; f++) {
                o << "      <face>";
                for (int c = 0; c < 3; c++) {
                    uint32_t id = f * 3 + c;
                    o << "<vref vertex=\"" << id << "\"";
                    if (hs.has_normal) o << " normal=\"" << id << "\"";
                    if (hs.has_texcoord) o << " texcoord=\"" << id << "\"";
                    o << "/>";
                }
                o << "</face>\n";
            }
            o << "    </mesh>\n";
        }
        o << "  </object>\n";
    }
    for (size_t li = 0; li < hs.lights.size(); li++) {
        const sray_light_desc &l = hs.lights[li];
        const char *ty = l.type == SRAY_LIGHT_POINT ? "point" : l.type == SRAY_LIGHT_SPHERE ? "sphere" : "box";
        o << "  <light type=\"" << ty << "\" color=\"" << vec3(l.color) << "\">\n";
        if (hs.has_tracks && li < hs.light_tracks.size()) write_tracks(o, hs.light_tracks[li], "    ");
        else o << "    <xform pos=\"" << vec3(l.pos) << "\"/>\n";
        if (l.type == SRAY_LIGHT_SPHERE) o << "    <sphere rad=\"" << f9(l.radius) << "\"/>\n";
        if (l.type == SRAY_LIGHT_BOX) {
            float mn[3] = {-0.5f * l.dim[0], -0.5f * l.dim[1], -0.5f * l.dim[2]};
            float mx[3] = {0.5f * l.dim[0], 0.5f * l.dim[1], 0.5f * l.dim[2]};
            o << "    <aabox min=\"" << vec3(mn) << "\" max=\"" << vec3(mx) << "\"/>\n";
        }
        o << "  </light>\n";
    }
    /* camera: written as a free camera with its position; generators use
     * axis-aligned free cameras or target cameras (target written below) */
    const sray_camera_desc &c = hs.camera;
    char fov[64];
    snprintf(fov, sizeof fov, "%.9g", c.vfov * 180.0 / M_PI);
    std::string fs = fov;
    if (fs.find('.') == std::string::npos) fs += ".0";
    bool ident = c.xform[0] == 1 && c.xform[5] == 1 && c.xform[10] == 1 && c.xform[1] == 0 && c.xform[2] == 0 &&
                 c.xform[4] == 0 && c.xform[6] == 0 && c.xform[8] == 0 && c.xform[9] == 0;
    float cp[3] = {c.xform[3], c.xform[7], c.xform[11]};
    const std::string shut = c.shutter > 0 ? " shutter=\"" + std::to_string(c.shutter) + "\"" : "";
    if (hs.has_tracks) {
        o << "  <camera type=\"" << (hs.cam_has_target ? "target" : "free") << "\" fov=\"" << fs << "\"" << shut
          << ">\n";
        write_tracks(o, hs.cam_tracks, "    ");
        if (hs.cam_has_target) {
            o << "    <target>\n";
            write_tracks(o, hs.cam_target_tracks, "      ");
            o << "    </target>\n";
        }
        o << "  </camera>\n";
    } else if (ident) {
        o << "  <camera type=\"free\" fov=\"" << fs << "\"" << shut << ">\n    <xform pos=\"" << vec3(cp)
          << "\"/>\n  </camera>\n";
    } else {
        float tg[3] = {cp[0] + c.xform[2], cp[1] + c.xform[6], cp[2] + c.xform[10]};
        if (hs.cam_has_target) {
            for (int k = 0; k < 3; k++) {
                cp[k] = hs.cam_pos[k];
                tg[k] = hs.cam_target[k];
            }
        }
        o << "  <camera type=\"target\" fov=\"" << fs << "\"" << shut << ">\n    <xform pos=\"" << vec3(cp) << "\"/>\n"
          << "    <target><xform pos=\"" << vec3(tg) << "\"/></target>\n  </camera>\n";
    }
    o << "</scene>\n";
    return o.str();
}

} // namespace sray

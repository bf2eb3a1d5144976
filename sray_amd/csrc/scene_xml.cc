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

/* xml_create_attr, src/xmltree.c:51-99: a value that starts with a digit
 * (or a sign and a digit) is a number — FLT if it contains '.', else INT —
 * and a VEC if it holds more than one space-separated number; unspecified
 * vector components are 1.0.  ".5" stays a string (Q16). */
XmlAttr xml_make_attr(const std::string &name, const std::string &val)
{
    XmlAttr a;
    a.name = name;
    a.str = val;
    const char *v = val.c_str();
    if (isdigit((unsigned char)v[0]) || ((v[0] == '+' || v[0] == '-') && isdigit((unsigned char)v[1]))) {
        a.type = strchr(v, '.') ? ATYPE_FLT : ATYPE_INT;
        a.ival = atoi(v);
        a.fval = (float)atof(v);
        a.vval[0] = a.vval[1] = a.vval[2] = a.vval[3] = 1.0f;
        int i;
        for (i = 0; i < 4; i++) {
            if (!*v) break;
            a.vval[i] = (float)atof(v);
            while (*v && !isspace((unsigned char)*v)) v++;
            while (*v && isspace((unsigned char)*v)) v++;
        }
        if (i > 1) a.type = ATYPE_VEC;
    }
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
/* transforms: XFormNode (src/anim.cc)                                */
/* ================================================================ */
namespace {

struct Quat {
    double s = 1, x = 0, y = 0, z = 0;
};

Quat quat_axis_angle(double ax, double ay, double az, double angle)
{
    /* Quaternion(axis, angle): s = cos(a/2), v = axis * sin(a/2) */
    Quat q;
    double h = angle / 2.0;
    q.s = cos(h);
    q.x = ax * sin(h);
    q.y = ay * sin(h);
    q.z = az * sin(h);
    return q;
}

void quat_to_mat(const Quat &q, double m[12])
{
    double s = q.s, x = q.x, y = q.y, z = q.z;
    mat_identity(m);
    m[0] = 1 - 2 * y * y - 2 * z * z;
    m[1] = 2 * x * y - 2 * s * z;
    m[2] = 2 * z * x + 2 * s * y;
    m[4] = 2 * x * y + 2 * s * z;
    m[5] = 1 - 2 * x * x - 2 * z * z;
    m[6] = 2 * y * z - 2 * s * x;
    m[8] = 2 * z * x - 2 * s * y;
    m[9] = 2 * y * z + 2 * s * x;
    m[10] = 1 - 2 * x * x - 2 * y * y;
}

template <class V> struct Key {
    int time;
    V val;
};

struct V3d {
    double x = 0, y = 0, z = 0;
};

/* XFormNode: position / rotation / scale tracks + pivot */
struct XForm {
    std::vector<Key<V3d>> pos, scale;
    std::vector<Key<Quat>> rot;
    V3d pivot;
    int interp_step = 0;

    template <class V> static void set_key(std::vector<Key<V>> &tr, int t, const V &v)
    {
        for (auto &k : tr)
            if (k.time == t) {
                k.val = v;
                return;
            }
        tr.push_back({t, v});
        std::sort(tr.begin(), tr.end(), [](const Key<V> &a, const Key<V> &b) { return a.time < b.time; });
    }

    /* XFormNode::load_xml, src/anim.cc:68-121 */
    void load(const XmlNode &n)
    {
        int msec = 0;
        const XmlAttr *a;
        if ((a = n.get("interpolate")) && a->str == "step") interp_step = 1;
        if ((a = n.get("time")) && a->type == ATYPE_INT) msec = a->ival;
        if ((a = n.get("pos")) && a->type == ATYPE_VEC) set_key(pos, msec, V3d{a->vval[0], a->vval[1], a->vval[2]});
        if ((a = n.get("rot")) && a->type == ATYPE_VEC) {
            Quat q;
            q.s = a->vval[0];
            q.x = a->vval[1];
            q.y = a->vval[2];
            q.z = a->vval[3];
            double l = sqrt(q.s * q.s + q.x * q.x + q.y * q.y + q.z * q.z);
            if (l > 0) {
                q.s /= l;
                q.x /= l;
                q.y /= l;
                q.z /= l;
            }
            set_key(rot, msec, q);
        }
        if ((a = n.get("axisrot")) && a->type == ATYPE_VEC) {
            double ang = a->vval[0] * M_PI / 180.0;
            set_key(rot, msec, quat_axis_angle(a->vval[1], a->vval[2], a->vval[3], ang));
        }
        if ((a = n.get("scale")) && a->type == ATYPE_VEC) set_key(scale, msec, V3d{a->vval[0], a->vval[1], a->vval[2]});
        if ((a = n.get("pivot")) && a->type == ATYPE_VEC) pivot = V3d{a->vval[0], a->vval[1], a->vval[2]};
    }

    /* Track<T>::operator(), src/anim.inl:188-261 (clamped, linear or step) */
    V3d eval_v(const std::vector<Key<V3d>> &tr, int t, V3d def) const
    {
        if (tr.empty()) return def;
        if (t <= tr.front().time) return tr.front().val;
        if (t >= tr.back().time) return tr.back().val;
        for (size_t i = 0; i + 1 < tr.size(); i++) {
            if (t >= tr[i].time && t < tr[i + 1].time) {
                if (interp_step) return tr[i].val;
                double u = (double)(t - tr[i].time) / (double)(tr[i + 1].time - tr[i].time);
                const V3d &a = tr[i].val, &b = tr[i + 1].val;
                return V3d{a.x + (b.x - a.x) * u, a.y + (b.y - a.y) * u, a.z + (b.z - a.z) * u};
            }
        }
        return tr.back().val;
    }
    Quat eval_q(int t) const
    {
        if (rot.empty()) return Quat();
        if (t <= rot.front().time) return rot.front().val;
        if (t >= rot.back().time) return rot.back().val;
        for (size_t i = 0; i + 1 < rot.size(); i++) {
            if (t >= rot[i].time && t < rot[i + 1].time) {
                if (interp_step) return rot[i].val;
                double u = (double)(t - rot[i].time) / (double)(rot[i + 1].time - rot[i].time);
                /* slerp */
                Quat a = rot[i].val, b = rot[i + 1].val;
                double c = a.s * b.s + a.x * b.x + a.y * b.y + a.z * b.z;
                if (c < 0) {
                    c = -c;
                    b.s = -b.s;
                    b.x = -b.x;
                    b.y = -b.y;
                    b.z = -b.z;
                }
                double wa, wb;
                if (c > 0.9995) {
                    wa = 1 - u;
                    wb = u;
                } else {
                    double th = acos(c), st = sin(th);
                    wa = sin((1 - u) * th) / st;
                    wb = sin(u * th) / st;
                }
                Quat q;
                q.s = a.s * wa + b.s * wb;
                q.x = a.x * wa + b.x * wb;
                q.y = a.y * wa + b.y * wb;
                q.z = a.z * wa + b.z * wb;
                return q;
            }
        }
        return rot.back().val;
    }
    V3d position(int t) const { return eval_v(pos, t, V3d{0, 0, 0}); }

    /* XFormNode::get_xform_matrix, src/anim.cc:391-419:
     * pivot * T * R * S * pivot^-1 */
    void matrix(int t, double m[12]) const
    {
        V3d p = position(t);
        V3d s = eval_v(scale, t, V3d{1, 1, 1});
        Quat q = eval_q(t);
        double P[12], T[12], R[12], S[12], NP[12], tmp[12];
        mat_identity(P);
        P[3] = pivot.x;
        P[7] = pivot.y;
        P[11] = pivot.z;
        mat_identity(NP);
        NP[3] = -pivot.x;
        NP[7] = -pivot.y;
        NP[11] = -pivot.z;
        mat_identity(T);
        T[3] = p.x;
        T[7] = p.y;
        T[11] = p.z;
        quat_to_mat(q, R);
        mat_identity(S);
        S[0] = s.x;
        S[5] = s.y;
        S[10] = s.z;
        mat_mul(P, T, tmp);
        mat_mul(tmp, R, tmp);
        mat_mul(tmp, S, tmp);
        mat_mul(tmp, NP, m);
    }
    /* two position keys and static rotation/scale -> linear translation */
    bool linear_motion(V3d &delta, int &t0, int &t1) const
    {
        if (pos.size() != 2 || rot.size() > 1 || scale.size() > 1 || interp_step) return false;
        delta = V3d{pos[1].val.x - pos[0].val.x, pos[1].val.y - pos[0].val.y, pos[1].val.z - pos[0].val.z};
        t0 = pos[0].time;
        t1 = pos[1].time;
        return true;
    }
};

void load_xforms(const XmlNode &n, XForm &xf)
{
    for (const XmlNode &c : n.chld)
        if (c.name == "xform") xf.load(c);
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

/* Material::load_xml + MatAttrib::load_xml, src/material.cc:36-124 */
bool load_material(const XmlNode &n, HostScene &hs, std::string &err)
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
        /* a "map" texture would need libimago; it fails to load and the
         * attribute keeps its plain colour, as the reference does when
         * Texture::load fails */
        const std::string &nm = an->str;
        if (nm == "diffuse") std::memcpy(m.diffuse, col, sizeof col);
        else if (nm == "specular") {
            std::memcpy(m.specular, col, sizeof col);
            m.flags |= SRAY_MAT_HAS_SPECULAR;
        } else if (nm == "shininess") m.shininess = col[0];
        else if (nm == "reflect") {
            m.reflect = col[0];
            m.flags |= SRAY_MAT_HAS_REFLECT;
        } else if (nm == "refract") {
            m.refract = col[0];
            m.flags |= SRAY_MAT_HAS_REFRACT;
        } else if (nm == "ior") m.ior = col[0];
    }
    hs.materials.push_back(m);
    hs.material_names.push_back(name);
    return true;
}

bool set_xform(sray_object_desc &o, const XForm &xf, int frame_time, std::string &err)
{
    double m[12], inv[12];
    V3d dlt;
    int t0, t1;
    if (xf.linear_motion(dlt, t0, t1)) {
        xf.matrix(t0, m);
        o.motion[0] = (float)dlt.x;
        o.motion[1] = (float)dlt.y;
        o.motion[2] = (float)dlt.z;
        o.motion_t0 = t0;
        o.motion_t1 = t1;
    } else {
        xf.matrix(frame_time, m);
    }
    if (!mat_inverse(m, inv)) {
        err = "singular object transform";
        return false;
    }
    mat_to_float(m, o.xform);
    mat_to_float(inv, o.inv_xform);
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
bool load_object(const XmlNode &n, HostScene &hs, int frame_time, std::string &err)
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
    XForm xf;
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
            xf.load(c);
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
    if (!set_xform(o, xf, frame_time, err)) return false;
    hs.objects.push_back(o);
    return true;
}

/* load_light + Light/SphLight/BoxLight::load_xml, src/light.cc:39-65,236-255,310-329 */
bool load_light(const XmlNode &n, HostScene &hs, int frame_time, std::string &err)
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
    XForm xf;
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
    V3d p = xf.position(frame_time);
    l.pos[0] = (float)p.x;
    l.pos[1] = (float)p.y;
    l.pos[2] = (float)p.z;
    double m[12];
    xf.matrix(0, m); /* BoxLight::get_point uses get_xform_matrix() at time 0 (Q13) */
    mat_to_float(m, l.xform);
    hs.lights.push_back(l);
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
        double R[12];
        quat_to_mat(quat_axis_angle(dx, dy, dz, roll), R);
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
bool load_camera(const XmlNode &n, HostScene &hs, int frame_time, std::string &err)
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
    XForm xf;
    load_xforms(n, xf);
    double m[12];
    if (!target) {
        xf.matrix(frame_time, m);
    } else {
        XForm tx;
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
        V3d pos = xf.position(frame_time), tg = tx.position(frame_time);
        double roll = tx.eval_v(tx.scale, frame_time, V3d{0, 0, 0}).x;
        double P[3] = {pos.x, pos.y, pos.z}, Tg[3] = {tg.x, tg.y, tg.z};
        target_camera_matrix(P, Tg, roll, m);
        hs.cam_has_target = true;
        for (int k = 0; k < 3; k++) {
            hs.cam_pos[k] = (float)P[k];
            hs.cam_target[k] = (float)Tg[k];
        }
    }
    mat_to_float(m, c.xform);
    hs.camera = c;
    return true;
}

} // namespace

/* Scene::load(FILE*), src/scene.cc:93-183: materials first, then env,
 * objects, lights and the (first) camera in document order */
bool load_scene_xml(const std::string &text, HostScene &hs, std::string &err)
{
    XmlNode root;
    if (!xml_parse(text, root, err)) return false;
    hs = HostScene();
    hs.camera = default_camera();
    int frame_time = 0;
    for (const XmlNode &n : root.chld)
        if (n.name == "material" && !load_material(n, hs, err)) return false;
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
            if (!load_object(n, hs, frame_time, err)) return false;
        } else if (n.name == "light") {
            if (!load_light(n, hs, frame_time, err)) return false;
        } else if (n.name == "camera") {
            HostScene tmp;
            if (!load_camera(n, hs.camera.max_dist != 0 && have_cam ? tmp : hs, frame_time, err)) return false;
            have_cam = true;
        }
    }
    hs.finish();
    return true;
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
        o << "    <mattr name=\"diffuse\" value=\"" << vec3(m.diffuse) << "\"/>\n";
        o << "    <mattr name=\"specular\" value=\"" << vec3(m.specular) << "\"/>\n";
        o << "    <mattr name=\"shininess\" value=\"" << f9(m.shininess) << "\"/>\n";
        o << "    <mattr name=\"reflect\" value=\"" << f9(m.reflect) << "\"/>\n";
        o << "    <mattr name=\"refract\" value=\"" << f9(m.refract) << "\"/>\n";
        o << "    <mattr name=\"ior\" value=\"" << f9(m.ior) << "\"/>\n";
        o << "  </material>\n";
    }
    for (const sray_object_desc &ob : hs.objects) {
        const char *ty = ob.type == SRAY_OBJ_SPHERE ? "sphere" : ob.type == SRAY_OBJ_CYLINDER ? "cylinder" : "mesh";
        o << "  <object type=\"" << ty << "\">\n";
        if (ob.material >= 0 && ob.material < (int)hs.material_names.size())
            o << "    <matref name=\"" << hs.material_names[ob.material] << "\"/>\n";
        /* generators only place objects by translation (+ optional motion) */
        float p0[3] = {ob.xform[3], ob.xform[7], ob.xform[11]};
        bool moving = (ob.motion[0] != 0 || ob.motion[1] != 0 || ob.motion[2] != 0) && ob.motion_t1 > ob.motion_t0;
        if (moving) {
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
    for (const sray_light_desc &l : hs.lights) {
        const char *ty = l.type == SRAY_LIGHT_POINT ? "point" : l.type == SRAY_LIGHT_SPHERE ? "sphere" : "box";
        o << "  <light type=\"" << ty << "\" color=\"" << vec3(l.color) << "\">\n";
        o << "    <xform pos=\"" << vec3(l.pos) << "\"/>\n";
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
    if (ident) {
        o << "  <camera type=\"free\" fov=\"" << fs << "\">\n    <xform pos=\"" << vec3(cp) << "\"/>\n  </camera>\n";
    } else {
        float tg[3] = {cp[0] + c.xform[2], cp[1] + c.xform[6], cp[2] + c.xform[10]};
        if (hs.cam_has_target) {
            for (int k = 0; k < 3; k++) {
                cp[k] = hs.cam_pos[k];
                tg[k] = hs.cam_target[k];
            }
        }
        o << "  <camera type=\"target\" fov=\"" << fs << "\">\n    <xform pos=\"" << vec3(cp) << "\"/>\n"
          << "    <target><xform pos=\"" << vec3(tg) << "\"/></target>\n  </camera>\n";
    }
    o << "</scene>\n";
    return o.str();
}

} // namespace sray

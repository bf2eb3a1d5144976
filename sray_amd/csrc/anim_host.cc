/*
 * anim_host.cc — keyframe tracks and transforms on the host (XFormNode).
 *
 * Restates Track<T>::operator() (src/anim.inl:188-261: repeat
 * extrapolation, the nearest-key interval of get_key_interval,
 * src/anim.inl:59-106, step / linear / cubic interpolation) and
 * XFormNode::get_xform_matrix (src/anim.cc:391-419: pivot * T * R * S *
 * pivot^-1), in double precision as the reference computes them.  The
 * renderer evaluates every animated transform at every millisecond of the
 * frame interval here, once per frame, and the device reads the rows at a
 * ray's time: object matrices and their inverses (Object::intersect at
 * ray.time, src/sphere.cc:88, src/mesh.cc:263), the camera matrix
 * (Camera::get_primary_ray, src/camera.cc:141) and light positions
 * (Light::get_point(tm), src/shader.cc:258, src/light.cc:106-109).
 *
 * Quaternion interpolation and the cubic spline come from libvmath, which
 * the reference does not vendor: slerp along the shorter arc (linear
 * weights for nearly equal keys) and the uniform Catmull-Rom spline are the
 * published forms.
 */
#include <cmath>
#include <cstring>

#include "host.h"

namespace sray {

namespace {

template <class K> int nearest_key(const std::vector<K> &keys, int start, int end, int time)
{
    /* Track::get_nearest_key (binary search), src/anim.inl:69-80 */
    if (start == end) return start;
    if (end - start == 1)
        return std::abs(time - keys[start].time) < std::abs(keys[end].time - time) ? start : end;
    int mid = (start + end) / 2;
    if (time < keys[mid].time) return nearest_key(keys, start, mid, time);
    if (time > keys[mid].time) return nearest_key(keys, mid + 1, end, time);
    return mid;
}

/* Track::operator() up to the interpolation: the extrapolated time and the
 * interval [start, end] (end = -1: the start key's value) */
template <class K> void key_interval(const std::vector<K> &keys, int extrap, int &time, int &start, int &end)
{
    if (keys.size() > 1) {
        int first = keys.front().time, last = keys.back().time;
        if (first != last && (time < first || time > last) && extrap == SR_EXTRAP_REPEAT) {
            while (time < first) time += last - first;
            time = (time - first) % (last - first) + first;
        }
    }
    int n = nearest_key(keys, 0, (int)keys.size() - 1, time);
    start = n;
    end = -1;
    if (time < keys[n].time && keys[n].time != keys.front().time) {
        start = n - 1;
        end = n;
    } else if (time > keys[n].time && keys[n].time != keys.back().time) {
        start = n;
        end = n + 1;
    }
}

void slerp(const double a[4], const double b0[4], double t, double out[4])
{
    double b[4] = {b0[0], b0[1], b0[2], b0[3]};
    double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    if (c < 0.0) {
        c = -c;
        for (int k = 0; k < 4; k++) b[k] = -b[k];
    }
    if (c > 1.0) c = 1.0;
    double wa = 1.0 - t, wb = t;
    const double th = std::acos(c), st = std::sin(th);
    if (st > 1e-9) {
        wa = std::sin((1.0 - t) * th) / st;
        wb = std::sin(t * th) / st;
    }
    for (int k = 0; k < 4; k++) out[k] = a[k] * wa + b[k] * wb;
}

void quat_matrix(const double q[4], double m[12])
{
    const double s = q[0], x = q[1], y = q[2], z = q[3];
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

} // namespace

void quat_from_axis_angle(double ax, double ay, double az, double angle, double q[4])
{
    /* Quaternion(axis, angle): s = cos(a/2), v = axis * sin(a/2) */
    const double h = angle / 2.0;
    q[0] = std::cos(h);
    q[1] = ax * std::sin(h);
    q[2] = ay * std::sin(h);
    q[3] = az * std::sin(h);
}

void quat_to_matrix(const double q[4], double m[12]) { quat_matrix(q, m); }

bool XFormTracks::vec(const std::vector<TrackKey3> &tr, int time, const double def[3], double out[3]) const
{
    if (tr.empty()) {
        std::memcpy(out, def, 3 * sizeof(double));
        return false;
    }
    int s, e;
    key_interval(tr, extrap, time, s, e);
    if (interp == SR_INTERP_STEP || e < 0) {
        std::memcpy(out, tr[s].v, 3 * sizeof(double));
        return true;
    }
    const double t = (double)(time - tr[s].time) / (double)(tr[e].time - tr[s].time);
    if (interp == SR_INTERP_CUBIC) {
        const double *v0 = tr[s > 0 ? s - 1 : s].v, *v1 = tr[s].v, *v2 = tr[e].v;
        const double *v3 = tr[e < (int)tr.size() - 1 ? e + 1 : e].v;
        const double t2 = t * t, t3 = t2 * t;
        for (int k = 0; k < 3; k++)
            out[k] = 0.5 * ((2.0 * v1[k]) + (-v0[k] + v2[k]) * t + (2.0 * v0[k] - 5.0 * v1[k] + 4.0 * v2[k] - v3[k]) * t2 +
                            (-v0[k] + 3.0 * v1[k] - 3.0 * v2[k] + v3[k]) * t3);
        return true;
    }
    for (int k = 0; k < 3; k++) out[k] = tr[s].v[k] + (tr[e].v[k] - tr[s].v[k]) * t;
    return true;
}

void XFormTracks::position(int time, double out[3]) const
{
    static const double zero[3] = {0, 0, 0};
    vec(pos, time, zero, out);
}

void XFormTracks::scaling(int time, double out[3]) const
{
    static const double one[3] = {1, 1, 1};
    vec(scale, time, one, out);
}

void XFormTracks::rotation(int time, double q[4]) const
{
    if (rot.empty()) {
        q[0] = 1;
        q[1] = q[2] = q[3] = 0;
        return;
    }
    int s, e;
    key_interval(rot, extrap, time, s, e);
    if (interp == SR_INTERP_STEP || e < 0) {
        std::memcpy(q, rot[s].q, 4 * sizeof(double));
        return;
    }
    /* linear and cubic both slerp (catmull_rom_spline on quaternions is a
     * slerp of the inner keys, src/anim.inl:270-275) */
    slerp(rot[s].q, rot[e].q, (double)(time - rot[s].time) / (double)(rot[e].time - rot[s].time), q);
}

void XFormTracks::matrix(int time, double m[12]) const
{
    double p[3], sc[3], q[4];
    position(time, p);
    scaling(time, sc);
    rotation(time, q);
    double P[12], T[12], R[12], S[12], NP[12], a[12], b[12];
    mat_identity(P);
    mat_identity(NP);
    mat_identity(T);
    mat_identity(S);
    for (int k = 0; k < 3; k++) {
        P[k * 4 + 3] = pivot[k];
        NP[k * 4 + 3] = -pivot[k];
        T[k * 4 + 3] = p[k];
    }
    S[0] = sc[0];
    S[5] = sc[1];
    S[10] = sc[2];
    quat_matrix(q, R);
    mat_mul(P, T, a);
    mat_mul(a, R, b);
    mat_mul(b, S, a);
    mat_mul(a, NP, m);
}

template <class K> static void put_key(std::vector<K> &tr, const K &k)
{
    /* Track::add_key: a key at an existing time replaces its value */
    for (K &x : tr)
        if (x.time == k.time) {
            x = k;
            return;
        }
    size_t i = 0;
    while (i < tr.size() && tr[i].time < k.time) i++;
    tr.insert(tr.begin() + (long)i, k);
}

void XFormTracks::set_pos(int t, double x, double y, double z) { put_key(pos, TrackKey3{t, {x, y, z}}); }
void XFormTracks::set_scale(int t, double x, double y, double z) { put_key(scale, TrackKey3{t, {x, y, z}}); }
void XFormTracks::set_rot(int t, const double q[4], const float raw[4], int axisrot)
{
    TrackKeyQ k;
    k.time = t;
    std::memcpy(k.q, q, sizeof k.q);
    std::memcpy(k.raw, raw, sizeof k.raw);
    k.axisrot = axisrot;
    put_key(rot, k);
}

bool XFormTracks::linear_translation() const
{
    bool no_pivot = pivot[0] == 0 && pivot[1] == 0 && pivot[2] == 0;
    return pos.size() == 2 && rot.size() <= 1 && scale.size() <= 1 && interp == SR_INTERP_LINEAR &&
           extrap != SR_EXTRAP_REPEAT && no_pivot;
}

/* ------------------------------------------------------------------ */
/* a frame: static matrices at msec and per-millisecond rows over the   */
/* frame interval (render(msec), src/render.cc:34-55)                   */
/* ------------------------------------------------------------------ */
bool HostScene::set_frame(int msec, std::string &err)
{
    if (!has_tracks) return true;
    const int shutter = camera.shutter > 0 ? camera.shutter : 0;
    frame_t0 = msec - shutter / 2;
    frame_t1 = frame_t0 + shutter;
    frame_msec = msec;
    const int n = frame_t1 - frame_t0 + 1;
    obj_tab.assign(objects.size(), std::vector<float>());
    for (size_t i = 0; i < objects.size() && i < obj_tracks.size(); i++) {
        const XFormTracks &xf = obj_tracks[i];
        sray_object_desc &o = objects[i];
        double m[12], inv[12];
        std::memset(o.motion, 0, sizeof o.motion);
        o.motion_t0 = o.motion_t1 = 0;
        if (xf.linear_translation()) {
            /* two position keys: a linear translation, evaluated on the
             * device (the same lerp, in float) */
            xf.matrix(xf.pos[0].time, m);
            for (int k = 0; k < 3; k++) o.motion[k] = (float)(xf.pos[1].v[k] - xf.pos[0].v[k]);
            o.motion_t0 = xf.pos[0].time;
            o.motion_t1 = xf.pos[1].time;
        } else {
            xf.matrix(msec, m);
        }
        if (!mat_inverse(m, inv)) {
            err = "singular object transform";
            return false;
        }
        mat_to_float(m, o.xform);
        mat_to_float(inv, o.inv_xform);
        if (xf.animated() && !xf.linear_translation()) {
            std::vector<float> &tab = obj_tab[i];
            tab.resize((size_t)n * 24);
            for (int t = 0; t < n; t++) {
                xf.matrix(frame_t0 + t, m);
                if (!mat_inverse(m, inv)) {
                    err = "singular object transform at " + std::to_string(frame_t0 + t) + " msec";
                    return false;
                }
                mat_to_float(m, &tab[(size_t)t * 24]);
                mat_to_float(inv, &tab[(size_t)t * 24 + 12]);
            }
        }
    }
    light_tab.assign(lights.size(), std::vector<float>());
    for (size_t i = 0; i < lights.size() && i < light_tracks.size(); i++) {
        const XFormTracks &xf = light_tracks[i];
        sray_light_desc &l = lights[i];
        double p[3], m[12];
        xf.position(msec, p);
        for (int k = 0; k < 3; k++) l.pos[k] = (float)p[k];
        xf.matrix(0, m); /* BoxLight::get_point: get_xform_matrix() at time 0 (Q13) */
        mat_to_float(m, l.xform);
        if (xf.pos.size() > 1 && l.type != SRAY_LIGHT_BOX) {
            std::vector<float> &tab = light_tab[i];
            tab.resize((size_t)n * 3);
            for (int t = 0; t < n; t++) {
                xf.position(frame_t0 + t, p);
                for (int k = 0; k < 3; k++) tab[(size_t)t * 3 + k] = (float)p[k];
            }
        }
    }
    cam_tab.clear();
    auto cam_matrix = [&](int t, double m[12]) {
        if (!cam_has_target) {
            cam_tracks.matrix(t, m);
            return;
        }
        double p[3], tg[3], roll[3];
        cam_tracks.position(t, p);
        cam_target_tracks.position(t, tg);
        const double zero[3] = {0, 0, 0};
        /* TargetCamera::get_roll: the target's scale track (src/camera.cc:205-214) */
        cam_target_tracks.vec(cam_target_tracks.scale, t, zero, roll);
        target_camera_matrix(p, tg, roll[0], m);
    };
    double m[12];
    cam_matrix(msec, m);
    mat_to_float(m, camera.xform);
    const bool cam_anim = cam_tracks.animated() || (cam_has_target && cam_target_tracks.animated());
    if (cam_anim) {
        cam_tab.resize((size_t)n * 12);
        for (int t = 0; t < n; t++) {
            cam_matrix(frame_t0 + t, m);
            mat_to_float(m, &cam_tab[(size_t)t * 12]);
        }
    }
    finish();
    return true;
}

} // namespace sray

/* Oracle: CPU restatement of the reference path tracer in plain C. TEST INFRASTRUCTURE ONLY;
 * see wf_oracle.h. Each function cites the reference file:line it restates (paths relative to
 * wavefront-vs-megakernel-path-tracing/). Build: gcc -O2 -std=c11 -ffp-contract=off (oracle/Makefile).
 */
#include "wf_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define K_PI 3.14159265358979323846f
#define K_INV_PI 0.31830988618379067154f
#define K_INV_4PI 0.07957747154594767f
#define K_INF 3.402823466e+38f
#define K_EPS 1e-4f

/* ---------------------------------------------------------------- math (include/wf/math.hpp) */
typedef struct { float x, y, z; } v3;
static inline v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 muls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 divs(v3 a, float s) { float r = 1.0f / s; return mk(a.x * r, a.y * r, a.z * r); } /* math.hpp:29 */
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len3(v3 a) { return sqrtf(dot3(a, a)); }
static inline v3 norm3(v3 a) { float l = len3(a); return l > 0.0f ? divs(a, l) : a; }
static inline float fmin_r(float a, float b) { return a < b ? a : b; }
static inline float fmax_r(float a, float b) { return a > b ? a : b; }
static inline float clamp_r(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline float lum3(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
static inline int finite1(float v) { return v - v == 0.0f; }
static inline v3 reflect3(v3 d, v3 n) { return sub(d, muls(n, 2.0f * dot3(d, n))); }
static void basis(v3 n, v3* t, v3* b) { /* math.hpp:54-60 */
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float c = n.x * n.y * a;
    *t = mk(1.0f + sign * n.x * n.x * a, sign * c, -sign * n.x);
    *b = mk(c, sign + n.y * n.y * a, -n.y);
}
/* Default: the C library's binary64 sin/cos rounded to binary32 (the correctly rounded values the
 * product's own sin/cos equals on every estimator angle). Built with OR_LIBM_F32_TRIG: the C
 * library's binary32 sinf/cosf, which is what the reference's host code calls
 * (include/wf/path_core.hpp:117-119,217-221), so that build reproduces the reference's renders bit
 * for bit. */
static inline void sincos_lib(float x, float* s, float* c) {
#ifdef OR_LIBM_F32_TRIG
    *s = sinf(x);
    *c = cosf(x);
#else
    *s = (float)sin((double)x);
    *c = (float)cos((double)x);
#endif
}

/* ---------------------------------------------------------------- rng (include/wf/rng.hpp) */
uint32_t or_hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU;
    x ^= x >> 15; x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
typedef struct { uint32_t k0, k1; } pkey;
static pkey path_key(uint32_t seed, uint32_t pixel, uint32_t sample) { /* rng.hpp:21-27 */
    uint32_t a = or_hash_u32(seed ^ 0x9e3779b9U);
    uint32_t b = or_hash_u32(seed + 0x85ebca6bU);
    pkey k;
    k.k0 = or_hash_u32(or_hash_u32(a + pixel * 0xc2b2ae35U) ^ (sample * 0x27d4eb2fU + 0x165667b1U));
    k.k1 = or_hash_u32(or_hash_u32(b ^ (pixel * 0x9e3779b1U)) + (sample * 0x85ebca77U + 0xc2b2ae3dU));
    return k;
}
static float rnd(pkey k, uint32_t depth, uint32_t dim) { /* rng.hpp:38-42 */
    uint32_t ctr = depth * 16u + dim;
    uint32_t h = or_hash_u32(k.k0 ^ or_hash_u32(ctr * 0x9e3779b9U + k.k1));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
void or_path_key(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t* k2) {
    pkey k = path_key(seed, pixel, sample);
    k2[0] = k.k0;
    k2[1] = k.k1;
}
float or_rng(uint32_t k0, uint32_t k1, uint32_t depth, uint32_t dim) {
    pkey k = {k0, k1};
    return rnd(k, depth, dim);
}
void or_sincos(const float* x, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) sincos_lib(x[i], &s[i], &c[i]);
}
enum { D_PX = 0, D_PY = 1, D_PICK = 2, D_LU = 3, D_LV = 4, D_BU = 5, D_BV = 6, D_RR = 7, D_DIEL = 8, D_FU = 9, D_FV = 10 };

/* ---------------------------------------------------------------- scene */
typedef struct { v3 c; float r; int mat; } sph;
typedef struct { int i0, i1, i2, mat; } tri;
typedef struct { int type; float ior, fuzz; v3 albedo, emission; } mtl;
typedef struct { v3 lo; int skip; v3 hi; int start, count; } bnode;
enum { LAMBERT = 0, METAL = 1, DIELECTRIC = 2, EMISSIVE = 3 };

struct or_scene {
    int ns, nv, nt, nm;
    sph* S;
    v3* V;
    tri* T;
    mtl* M;
    v3 env, sky_lo, sky_hi, sun_dir, sun_E;
    int sky_mode, sky_light, sun_light;
    v3 cam_pos, cam_at, cam_up;
    float vfov;
    int W, H;
    v3 rig_o, rig_ll, rig_h, rig_v;
    int nn;
    bnode* N;
    int* order;
    int nl;
    int* lights;
    int n_sel;
};

static int ibits(float f) { int v; memcpy(&v, &f, 4); return v; }
static int emits(const mtl* m) { return m->emission.x > 0.0f || m->emission.y > 0.0f || m->emission.z > 0.0f; }

/* ---- BVH (src/bvh_build.cpp) ---- */
typedef struct { v3 lo, hi; } box;
static box box_empty(void) { box b; b.lo = mk(K_INF, K_INF, K_INF); b.hi = mk(-K_INF, -K_INF, -K_INF); return b; }
static void grow_pt(box* b, v3 p) { /* bvh_build.cpp:14-17 */
    b->lo.x = fmin_r(b->lo.x, p.x); b->lo.y = fmin_r(b->lo.y, p.y); b->lo.z = fmin_r(b->lo.z, p.z);
    b->hi.x = fmax_r(b->hi.x, p.x); b->hi.y = fmax_r(b->hi.y, p.y); b->hi.z = fmax_r(b->hi.z, p.z);
}
static void grow_box(box* b, box o) {
    b->lo.x = fmin_r(b->lo.x, o.lo.x); b->lo.y = fmin_r(b->lo.y, o.lo.y); b->lo.z = fmin_r(b->lo.z, o.lo.z);
    b->hi.x = fmax_r(b->hi.x, o.hi.x); b->hi.y = fmax_r(b->hi.y, o.hi.y); b->hi.z = fmax_r(b->hi.z, o.hi.z);
}
static float box_area(box b) { v3 d = sub(b.hi, b.lo); return 2.0f * (d.x * d.y + d.y * d.z + d.z * d.x); }
static box prim_box(const or_scene* s, int p) { /* bvh_build.cpp:28-39 */
    box b;
    if (p < s->ns) {
        v3 r = mk(s->S[p].r, s->S[p].r, s->S[p].r);
        b.lo = sub(s->S[p].c, r);
        b.hi = add(s->S[p].c, r);
        return b;
    }
    const tri* t = &s->T[p - s->ns];
    b.lo = b.hi = s->V[t->i0];
    grow_pt(&b, s->V[t->i1]);
    grow_pt(&b, s->V[t->i2]);
    return b;
}
static v3 prim_cen(const or_scene* s, int p) { /* bvh_build.cpp:41-48 */
    if (p < s->ns) return s->S[p].c;
    const tri* t = &s->T[p - s->ns];
    v3 c = add(add(s->V[t->i0], s->V[t->i1]), s->V[t->i2]);
    return muls(c, 1.0f / 3.0f);
}
static float axis_v(v3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

typedef struct { v3 lo, hi; int left, right, start, count, size; } tnode;
typedef struct { const or_scene* s; tnode* tn; int ntn, captn; int* order; int norder; } tbuild;

static int tpush(tbuild* B) {
    if (B->ntn == B->captn) {
        B->captn = B->captn ? 2 * B->captn : 1024;
        B->tn = (tnode*)realloc(B->tn, sizeof(tnode) * (size_t)B->captn);
    }
    return B->ntn++;
}

/* build_temp_tree, bvh_build.cpp:60-187 */
static int build_tmp(tbuild* B, const int* prims, int n) {
    if (n <= 0) return -1;
    const or_scene* s = B->s;
    int id = tpush(B);
    tnode node;
    memset(&node, 0, sizeof(node));
    node.left = node.right = -1;
    node.size = 1;
    box bounds = box_empty(), cb = box_empty();
    for (int k = 0; k < n; ++k) {
        grow_box(&bounds, prim_box(s, prims[k]));
        grow_pt(&cb, prim_cen(s, prims[k]));
    }
    node.lo = bounds.lo;
    node.hi = bounds.hi;
    if (n <= 4) {
        node.start = B->norder;
        node.count = n;
        for (int k = 0; k < n; ++k) B->order[B->norder++] = prims[k];
        B->tn[id] = node;
        return id;
    }
    node.count = 0;
    node.start = -1;
    v3 ext = sub(cb.hi, cb.lo);
    int axis = 0;
    if (ext.y > ext.x) axis = 1;
    if (ext.z > ext.x && ext.z > ext.y) axis = 2;
    int* L = (int*)malloc(sizeof(int) * (size_t)n);
    int* R = (int*)malloc(sizeof(int) * (size_t)n);
    int nl = 0, nr = 0;
    if (ext.x < 1e-6f && ext.y < 1e-6f && ext.z < 1e-6f) {
        int mid = n / 2;
        for (int k = 0; k < mid; ++k) L[nl++] = prims[k];
        for (int k = mid; k < n; ++k) R[nr++] = prims[k];
    } else {
        float amin = axis_v(cb.lo, axis), amax = axis_v(cb.hi, axis), range = amax - amin;
        box bins[8];
        int cnt[8];
        for (int i = 0; i < 8; ++i) { bins[i] = box_empty(); cnt[i] = 0; }
        for (int k = 0; k < n; ++k) {
            int bi = (int)((axis_v(prim_cen(s, prims[k]), axis) - amin) / range * (8 - 0.0001f));
            bi = bi < 0 ? 0 : (bi >= 8 ? 7 : bi);
            grow_box(&bins[bi], prim_box(s, prims[k]));
            cnt[bi]++;
        }
        float best_cost = 1e9f;
        int best = 1;
        for (int split = 1; split < 8; ++split) {
            box lb = box_empty(), rb = box_empty();
            int lc = 0, rc = 0;
            for (int i = 0; i < split; ++i) { grow_box(&lb, bins[i]); lc += cnt[i]; }
            for (int i = split; i < 8; ++i) { grow_box(&rb, bins[i]); rc += cnt[i]; }
            float cost = (lc * box_area(lb) + rc * box_area(rb)) / box_area(bounds);
            if (cost < best_cost) { best_cost = cost; best = split; }
        }
        for (int k = 0; k < n; ++k) {
            int bi = (int)((axis_v(prim_cen(s, prims[k]), axis) - amin) / range * (8 - 0.0001f));
            bi = bi < 0 ? 0 : (bi >= 8 ? 7 : bi);
            if (bi < best) L[nl++] = prims[k];
            else R[nr++] = prims[k];
        }
        if (nl == 0 || nr == 0) {
            int mid = n / 2;
            nl = nr = 0;
            for (int k = 0; k < mid; ++k) L[nl++] = prims[k];
            for (int k = mid; k < n; ++k) R[nr++] = prims[k];
        }
    }
    B->tn[id] = node;
    int l = build_tmp(B, L, nl);
    free(L);
    int r = build_tmp(B, R, nr);
    free(R);
    if (l >= 0) { B->tn[id].left = l; B->tn[id].size += B->tn[l].size; }
    if (r >= 0) { B->tn[id].right = r; B->tn[id].size += B->tn[r].size; }
    return id;
}

/* flatten, bvh_build.cpp:191-226 */
static int flatten(const tbuild* B, int t, int skip, bnode* out, int* nout) {
    if (t < 0) return 0;
    const tnode* tn = &B->tn[t];
    int idx = (*nout)++;
    out[idx].lo = tn->lo;
    out[idx].hi = tn->hi;
    out[idx].skip = skip;
    out[idx].start = tn->start;
    out[idx].count = tn->count;
    if (tn->count > 0) return 1;
    int right_idx = idx + 1 + (tn->left >= 0 ? B->tn[tn->left].size : 0);
    int c = 1;
    if (tn->left >= 0) c += flatten(B, tn->left, right_idx, out, nout);
    if (tn->right >= 0) c += flatten(B, tn->right, skip, out, nout);
    return c;
}

static void camera_rig(or_scene* s) { /* include/wf/scene.hpp:129-141 */
    float aspect = (float)s->W / (float)s->H;
    float h = tanf(s->vfov * 0.5f * K_PI / 180.0f);
    v3 w = norm3(sub(s->cam_pos, s->cam_at));
    v3 u = norm3(cross3(s->cam_up, w));
    v3 v = cross3(w, u);
    s->rig_o = s->cam_pos;
    s->rig_h = muls(u, 2.0f * h * aspect);
    s->rig_v = muls(v, 2.0f * h);
    s->rig_ll = sub(sub(sub(s->rig_o, muls(s->rig_h, 0.5f)), muls(s->rig_v, 0.5f)), w);
}

or_scene* or_scene_new(const float* spheres, int ns, const float* vertices, int nv, const int* triangles, int nt,
                       const float* materials, int nm, const float* env, const float* cam, int width, int height) {
    if (ns + nt <= 0 || nm <= 0 || width < 1 || height < 1) return NULL;
    or_scene* s = (or_scene*)calloc(1, sizeof(or_scene));
    s->ns = ns; s->nv = nv; s->nt = nt; s->nm = nm; s->W = width; s->H = height;
    s->S = (sph*)malloc(sizeof(sph) * (size_t)(ns > 0 ? ns : 1));
    s->V = (v3*)malloc(sizeof(v3) * (size_t)(nv > 0 ? nv : 1));
    s->T = (tri*)malloc(sizeof(tri) * (size_t)(nt > 0 ? nt : 1));
    s->M = (mtl*)malloc(sizeof(mtl) * (size_t)nm);
    for (int i = 0; i < ns; ++i) {
        const float* p = spheres + 8 * i;
        s->S[i].c = mk(p[0], p[1], p[2]);
        s->S[i].r = p[3];
        s->S[i].mat = ibits(p[4]);
    }
    for (int i = 0; i < nv; ++i) s->V[i] = mk(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
    for (int i = 0; i < nt; ++i) {
        s->T[i].i0 = triangles[4 * i]; s->T[i].i1 = triangles[4 * i + 1];
        s->T[i].i2 = triangles[4 * i + 2]; s->T[i].mat = triangles[4 * i + 3];
    }
    for (int i = 0; i < nm; ++i) {
        const float* m = materials + 12 * i;
        s->M[i].type = ibits(m[0]);
        s->M[i].ior = m[1];
        s->M[i].fuzz = m[2];
        s->M[i].albedo = mk(m[4], m[5], m[6]);
        s->M[i].emission = mk(m[8], m[9], m[10]);
    }
    s->env = mk(env[0], env[1], env[2]);
    s->sky_mode = ibits(env[3]);
    s->sky_lo = mk(env[4], env[5], env[6]);
    s->sky_light = ibits(env[7]);
    s->sky_hi = mk(env[8], env[9], env[10]);
    s->sun_light = ibits(env[11]);
    s->sun_dir = mk(env[12], env[13], env[14]);
    s->sun_E = mk(env[16], env[17], env[18]);
    s->cam_pos = mk(cam[0], cam[1], cam[2]);
    s->cam_at = mk(cam[3], cam[4], cam[5]);
    s->cam_up = mk(cam[6], cam[7], cam[8]);
    s->vfov = cam[9];
    camera_rig(s);
    /* lights: emissive spheres, then emissive triangles (src/scene.cpp:68-91) */
    s->lights = (int*)malloc(sizeof(int) * (size_t)(ns + nt));
    for (int i = 0; i < ns; ++i)
        if (emits(&s->M[s->S[i].mat])) s->lights[s->nl++] = i;
    for (int i = 0; i < nt; ++i)
        if (emits(&s->M[s->T[i].mat])) s->lights[s->nl++] = ns + i;
    s->n_sel = s->nl + (s->sky_light ? 1 : 0) + (s->sun_light ? 1 : 0);
    /* BVH (src/bvh_build.cpp:228-251) */
    int np = ns + nt;
    int* prims = (int*)malloc(sizeof(int) * (size_t)np);
    for (int i = 0; i < np; ++i) prims[i] = i;
    tbuild B;
    memset(&B, 0, sizeof(B));
    B.s = s;
    B.order = (int*)malloc(sizeof(int) * (size_t)np);
    int root = build_tmp(&B, prims, np);
    free(prims);
    s->N = (bnode*)malloc(sizeof(bnode) * (size_t)B.ntn);
    s->nn = 0;
    flatten(&B, root, -1, s->N, &s->nn);
    s->order = B.order;
    free(B.tn);
    return s;
}

void or_scene_free(or_scene* s) {
    if (!s) return;
    free(s->S); free(s->V); free(s->T); free(s->M); free(s->N); free(s->order); free(s->lights);
    free(s);
}
int or_scene_n_nodes(const or_scene* s) { return s->nn; }
int or_scene_n_lights(const or_scene* s) { return s->nl; }

void or_scene_bvh(const or_scene* s, int* nodes, int* prim_order) {
    if (nodes)
        for (int i = 0; i < s->nn; ++i) {
            int* p = nodes + 12 * i;
            const bnode* n = &s->N[i];
            p[0] = ibits(n->lo.x); p[1] = ibits(n->lo.y); p[2] = ibits(n->lo.z); p[3] = n->skip;
            p[4] = ibits(n->hi.x); p[5] = ibits(n->hi.y); p[6] = ibits(n->hi.z);
            p[7] = n->start; p[8] = n->count; p[9] = p[10] = p[11] = 0;
        }
    if (prim_order) memcpy(prim_order, s->order, sizeof(int) * (size_t)(s->ns + s->nt));
}

void or_scene_set_view(or_scene* s, const float* cam, int width, int height) {
    s->cam_pos = mk(cam[0], cam[1], cam[2]);
    s->cam_at = mk(cam[3], cam[4], cam[5]);
    s->cam_up = mk(cam[6], cam[7], cam[8]);
    s->vfov = cam[9];
    s->W = width;
    s->H = height;
    camera_rig(s);
}
void or_camera_rig(const or_scene* s, float* r) {
    v3 v[4] = {s->rig_o, s->rig_ll, s->rig_h, s->rig_v};
    for (int i = 0; i < 4; ++i) { r[3 * i] = v[i].x; r[3 * i + 1] = v[i].y; r[3 * i + 2] = v[i].z; }
}

/* FNV-1a 64 over the canonical scene bytes (src/scene.cpp:107-160); extension fields appended
 * only when any is set. */
typedef struct { uint64_t h; } fnv;
static void fput(fnv* f, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { f->h ^= b[i]; f->h *= 1099511628211ULL; }
}
uint64_t or_scene_hash(const or_scene* s) {
    fnv f = {14695981039346656037ULL};
    for (int i = 0; i < s->ns; ++i) { fput(&f, &s->S[i].c, 12); fput(&f, &s->S[i].r, 4); fput(&f, &s->S[i].mat, 4); }
    for (int i = 0; i < s->nv; ++i) fput(&f, &s->V[i], 12);
    for (int i = 0; i < s->nt; ++i) fput(&f, &s->T[i], 16);
    for (int i = 0; i < s->nm; ++i) {
        fput(&f, &s->M[i].type, 4); fput(&f, &s->M[i].ior, 4); fput(&f, &s->M[i].albedo, 12); fput(&f, &s->M[i].emission, 12);
    }
    fput(&f, &s->env, 12);
    fput(&f, &s->cam_pos, 12); fput(&f, &s->cam_at, 12); fput(&f, &s->cam_up, 12); fput(&f, &s->vfov, 4);
    fput(&f, &s->W, 4); fput(&f, &s->H, 4);
    int ext = s->sky_mode != 0 || s->sky_light || s->sun_light;
    for (int i = 0; i < s->nm; ++i) ext = ext || s->M[i].fuzz != 0.0f;
    if (ext) {
        for (int i = 0; i < s->nm; ++i) fput(&f, &s->M[i].fuzz, 4);
        fput(&f, &s->sky_mode, 4); fput(&f, &s->sky_lo, 12); fput(&f, &s->sky_hi, 12); fput(&f, &s->sky_light, 4);
        fput(&f, &s->sun_light, 4); fput(&f, &s->sun_dir, 12); fput(&f, &s->sun_E, 12);
    }
    return f.h;
}

/* ---------------------------------------------------------------- intersection (include/wf/bvh.hpp) */
typedef struct { v3 o; float tmin; v3 d; float tmax; } ray;

static int hit_sphere(const sph* sp, const ray* r, float* t_out) { /* bvh.hpp:25-41 */
    v3 oc = sub(r->o, sp->c);
    float a = dot3(r->d, r->d);
    float b = dot3(oc, r->d);
    v3 w = sub(oc, muls(r->d, b / a));
    float disc = sp->r * sp->r - dot3(w, w);
    if (disc < 0.0f) return 0;
    float mid = -b / a;
    float sq = sqrtf(disc / a);
    float t = mid - sq;
    if (!(t > r->tmin && t < r->tmax)) {
        t = mid + sq;
        if (!(t > r->tmin && t < r->tmax)) return 0;
    }
    *t_out = t;
    return 1;
}
static int hit_tri(v3 p0, v3 p1, v3 p2, const ray* r, float* t_out) { /* bvh.hpp:44-61 */
    v3 e1 = sub(p1, p0), e2 = sub(p2, p0);
    v3 pv = cross3(r->d, e2);
    float det = dot3(e1, pv);
    if (det == 0.0f) return 0;
    float inv = 1.0f / det;
    v3 tv = sub(r->o, p0);
    float u = dot3(tv, pv) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    v3 qv = cross3(tv, e1);
    float v = dot3(r->d, qv) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = dot3(e2, qv) * inv;
    if (!(t > r->tmin && t < r->tmax)) return 0;
    *t_out = t;
    return 1;
}
static int hit_box(v3 lo, v3 hi, v3 o, v3 inv, float tmin, float tmax) { /* bvh.hpp:64-78 */
    float t0 = tmin, t1 = tmax, tn, tf, sw;
    tn = (lo.x - o.x) * inv.x; tf = (hi.x - o.x) * inv.x;
    if (tn > tf) { sw = tn; tn = tf; tf = sw; }
    t0 = tn > t0 ? tn : t0; t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return 0;
    tn = (lo.y - o.y) * inv.y; tf = (hi.y - o.y) * inv.y;
    if (tn > tf) { sw = tn; tn = tf; tf = sw; }
    t0 = tn > t0 ? tn : t0; t1 = tf < t1 ? tf : t1;
    if (t0 > t1) return 0;
    tn = (lo.z - o.z) * inv.z; tf = (hi.z - o.z) * inv.z;
    if (tn > tf) { sw = tn; tn = tf; tf = sw; }
    t0 = tn > t0 ? tn : t0; t1 = tf < t1 ? tf : t1;
    return t0 <= t1;
}
static int hit_prim(const or_scene* s, int p, const ray* r, float* t) { /* bvh.hpp:80-87 */
    if (p < s->ns) return hit_sphere(&s->S[p], r, t);
    const tri* tr = &s->T[p - s->ns];
    return hit_tri(s->V[tr->i0], s->V[tr->i1], s->V[tr->i2], r, t);
}
typedef struct { float t; int prim; } hitrec;
static int intersect(const or_scene* s, ray r, hitrec* h, int anyhit) { /* bvh.hpp:91-116 */
    h->t = r.tmax;
    h->prim = -1;
    v3 inv = mk(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    int n = 0;
    while (n >= 0 && n < s->nn) {
        const bnode* nd = &s->N[n];
        if (hit_box(nd->lo, nd->hi, r.o, inv, r.tmin, r.tmax)) {
            if (nd->count > 0) {
                for (int i = 0; i < nd->count; ++i) {
                    int p = s->order[nd->start + i];
                    float t;
                    if (hit_prim(s, p, &r, &t)) {
                        r.tmax = t; h->t = t; h->prim = p;
                        if (anyhit) return 1;
                    }
                }
                n = nd->skip;
            } else {
                n = n + 1;
            }
        } else {
            n = nd->skip;
        }
    }
    return h->prim >= 0;
}
void or_intersect(const or_scene* s, const float* rays, int n, int anyhit, float* t, int* prim) {
    for (int i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        ray r = {mk(q[0], q[1], q[2]), q[3], mk(q[4], q[5], q[6]), q[7]};
        hitrec h;
        intersect(s, r, &h, anyhit);
        t[i] = h.t;
        prim[i] = h.prim;
    }
}
void or_intersect_brute(const or_scene* s, const float* rays, int n, float* t, int* prim) { /* bvh.hpp:119-129 */
    for (int i = 0; i < n; ++i) {
        const float* q = rays + 8 * i;
        ray r = {mk(q[0], q[1], q[2]), q[3], mk(q[4], q[5], q[6]), q[7]};
        hitrec h = {r.tmax, -1};
        for (int p = 0; p < s->ns + s->nt; ++p) {
            float tt;
            if (hit_prim(s, p, &r, &tt)) { r.tmax = tt; h.t = tt; h.prim = p; }
        }
        t[i] = h.t;
        prim[i] = h.prim;
    }
}

/* ---------------------------------------------------------------- estimator (include/wf/path_core.hpp) */
typedef struct {
    ray r;
    v3 T, L;
    int depth, specular;
    float prev_pdf;
    pkey key;
} pstate;
typedef struct { v3 p, n, n_out; int inside, material; } surf;

static float power_h(float a, float b) { float a2 = a * a, b2 = b * b, s = a2 + b2; return s > 0.0f ? a2 / s : 0.0f; }
static float prim_area(const or_scene* s, int p) { /* path_core.hpp:44-52 */
    if (p < s->ns) { float r = s->S[p].r; return 4.0f * K_PI * r * r; }
    const tri* t = &s->T[p - s->ns];
    return 0.5f * len3(cross3(sub(s->V[t->i1], s->V[t->i0]), sub(s->V[t->i2], s->V[t->i0])));
}
static v3 sky(const or_scene* s, v3 d) {
    if (s->sky_mode == 0) return s->env;
    float t = 0.5f * (d.y + 1.0f);
    return add(muls(s->sky_lo, 1.0f - t), muls(s->sky_hi, t));
}
static v3 sph_dir(float u1, float u2) {
    float z = 1.0f - 2.0f * u1;
    float r = sqrtf(fmax_r(0.0f, 1.0f - z * z));
    float phi = 2.0f * K_PI * u2, sp, cp;
    sincos_lib(phi, &sp, &cp);
    return mk(r * cp, r * sp, z);
}
static pstate init_path(const or_scene* s, int px, int py, uint32_t sample, uint32_t seed) { /* path_core.hpp:59-82 */
    pstate ps;
    ps.key = path_key(seed, (uint32_t)(py * s->W + px), sample);
    float jx = rnd(ps.key, 0, D_PX), jy = rnd(ps.key, 0, D_PY);
    float u = ((float)px + jx) / (float)s->W;
    float v = ((float)(s->H - 1 - py) + jy) / (float)s->H;
    v3 target = add(add(s->rig_ll, muls(s->rig_h, u)), muls(s->rig_v, v));
    ps.r.o = s->rig_o;
    ps.r.d = norm3(sub(target, s->rig_o));
    ps.r.tmin = K_EPS;
    ps.r.tmax = K_INF;
    ps.T = mk(1.0f, 1.0f, 1.0f);
    ps.L = mk(0.0f, 0.0f, 0.0f);
    ps.depth = 0;
    ps.specular = 1;
    ps.prev_pdf = 0.0f;
    return ps;
}
static surf make_surf(const or_scene* s, const ray* r, const hitrec* h) { /* path_core.hpp:84-101 */
    surf f;
    f.p = add(r->o, muls(r->d, h->t));
    if (h->prim < s->ns) {
        const sph* sp = &s->S[h->prim];
        f.n_out = divs(sub(f.p, sp->c), sp->r);
        f.material = sp->mat;
    } else {
        const tri* t = &s->T[h->prim - s->ns];
        f.n_out = norm3(cross3(sub(s->V[t->i1], s->V[t->i0]), sub(s->V[t->i2], s->V[t->i0])));
        f.material = t->mat;
    }
    f.inside = dot3(f.n_out, r->d) > 0.0f ? 1 : 0;
    f.n = f.inside ? neg(f.n_out) : f.n_out;
    return f;
}
typedef struct { v3 p, n; float pdf; v3 Le; int kind; } lsample;
static lsample sample_light(const or_scene* s, v3 from, pkey key, uint32_t depth) { /* path_core.hpp:104-140 */
    lsample ls;
    ls.pdf = 0.0f; ls.p = from; ls.n = mk(0, 0, 1); ls.Le = mk(0, 0, 0); ls.kind = 0;
    if (s->n_sel <= 0) return ls;
    int idx = (int)(rnd(key, depth, D_PICK) * (float)s->n_sel);
    if (idx >= s->n_sel) idx = s->n_sel - 1;
    float u1 = rnd(key, depth, D_LU), u2 = rnd(key, depth, D_LV);
    if (idx >= s->nl) {
        if (s->sky_light && idx == s->nl) {
            ls.kind = 1;
            ls.n = sph_dir(u1, u2);
            ls.Le = sky(s, ls.n);
            ls.pdf = K_INV_4PI / (float)s->n_sel;
        } else {
            ls.kind = 2;
            ls.n = s->sun_dir;
            ls.Le = s->sun_E;
            ls.pdf = 1.0f / (float)s->n_sel;
        }
        return ls;
    }
    int prim = s->lights[idx];
    float area = prim_area(s, prim);
    if (prim < s->ns) {
        const sph* sp = &s->S[prim];
        v3 dir = sph_dir(u1, u2);
        ls.n = dir;
        ls.p = add(sp->c, muls(dir, sp->r));
        ls.Le = s->M[sp->mat].emission;
    } else {
        const tri* t = &s->T[prim - s->ns];
        float su = sqrtf(u1);
        float b0 = 1.0f - su, b1 = u2 * su;
        v3 p0 = s->V[t->i0], p1 = s->V[t->i1], p2 = s->V[t->i2];
        ls.p = add(add(muls(p0, 1.0f - b0 - b1), muls(p1, b0)), muls(p2, b1));
        ls.n = norm3(cross3(sub(p1, p0), sub(p2, p0)));
        ls.Le = s->M[t->mat].emission;
    }
    v3 d = sub(ls.p, from);
    float dist2 = dot3(d, d);
    float dist = sqrtf(dist2);
    if (dist <= 0.0f) return ls;
    float cosl = fabsf(dot3(ls.n, d)) / dist;
    if (cosl <= 1e-6f) return ls;
    ls.pdf = (dist2 / (cosl * area)) / (float)s->n_sel;
    return ls;
}
static int hit_valid(const or_scene* s, const hitrec* h) { /* src/gpu/engine.cuh:44-48 */
    if (h->prim < 0) return 1;
    return h->prim < s->ns + s->nt && finite1(h->t) && h->t > 0.0f;
}
static int classify(const or_scene* s, pstate* ps, const hitrec* h, int max_depth, surf* f) { /* path_core.hpp:151-170 */
    if (h->prim < 0) {
        if (s->sky_light && !ps->specular && ps->prev_pdf > 0.0f) {
            float w = power_h(ps->prev_pdf, K_INV_4PI / (float)s->n_sel);
            ps->L = add(ps->L, muls(mulv(ps->T, sky(s, ps->r.d)), w));
        } else {
            ps->L = add(ps->L, mulv(ps->T, sky(s, ps->r.d)));
        }
        return -1;
    }
    *f = make_surf(s, &ps->r, h);
    const mtl* m = &s->M[f->material];
    if (emits(m)) {
        float w = 1.0f;
        if (!ps->specular && ps->prev_pdf > 0.0f) {
            float cosl = fabsf(dot3(f->n_out, ps->r.d));
            float pl = cosl <= 1e-6f ? 0.0f : ((h->t * h->t) / (cosl * prim_area(s, h->prim))) / (float)s->n_sel;
            w = power_h(ps->prev_pdf, pl);
        }
        ps->L = add(ps->L, muls(mulv(ps->T, m->emission), w));
    }
    if (m->type == EMISSIVE) return -1;
    if (ps->depth >= max_depth) return -1;
    return m->type;
}
static int roulette(pstate* ps) { /* path_core.hpp:173-179 */
    if (ps->depth < 3) return 1;
    float q = clamp_r(fmax_r(ps->T.x, fmax_r(ps->T.y, ps->T.z)), 0.05f, 0.95f);
    if (rnd(ps->key, (uint32_t)ps->depth, D_RR) >= q) return 0;
    ps->T = divs(ps->T, q);
    return 1;
}
static void next_ray(pstate* ps, v3 o, v3 d) { ps->r.o = o; ps->r.d = d; ps->r.tmin = K_EPS; ps->r.tmax = K_INF; }
static int shade_lambert(const or_scene* s, pstate* ps, const surf* f, ray* sh, v3* contrib, int* has) { /* path_core.hpp:192-228 */
    const mtl* m = &s->M[f->material];
    uint32_t d = (uint32_t)ps->depth;
    *has = 0;
    v3 org = add(f->p, muls(f->n, K_EPS));
    if (s->n_sel > 0) {
        lsample ls = sample_light(s, org, ps->key, d);
        if (ls.pdf > 0.0f) {
            v3 wi;
            float tmax;
            if (ls.kind == 0) {
                v3 to = sub(ls.p, org);
                float dist = len3(to);
                wi = divs(to, dist);
                tmax = dist * (1.0f - 1e-3f);
            } else {
                wi = ls.n;
                tmax = K_INF;
            }
            float cs = dot3(f->n, wi);
            if (cs > 0.0f) {
                float pb = cs * K_INV_PI;
                float w = ls.kind == 2 ? 1.0f : power_h(ls.pdf, pb);
                *contrib = muls(mulv(mulv(ps->T, m->albedo), ls.Le), cs * K_INV_PI * w / ls.pdf);
                sh->o = org; sh->d = wi; sh->tmin = K_EPS; sh->tmax = tmax;
                *has = 1;
            }
        }
    }
    float u1 = rnd(ps->key, d, D_BU), u2 = rnd(ps->key, d, D_BV);
    float r = sqrtf(u1), phi = 2.0f * K_PI * u2, sp, cp;
    sincos_lib(phi, &sp, &cp);
    v3 t, b;
    basis(f->n, &t, &b);
    float z = sqrtf(fmax_r(0.0f, 1.0f - u1));
    v3 dir = norm3(add(add(muls(t, r * cp), muls(b, r * sp)), muls(f->n, z)));
    ps->T = mulv(ps->T, m->albedo);
    ps->specular = 0;
    ps->prev_pdf = dot3(dir, f->n) * K_INV_PI;
    ps->depth += 1;
    next_ray(ps, org, dir);
    return roulette(ps);
}
/* Mirror direction about the facing normal n, perturbed by fuzz times a uniform unit vector
 * (dims 9-10) when fuzz > 0; 0 if it points below the surface (the path is absorbed). */
static int metal_dir(v3 din, v3 n, float fuzz, pkey key, uint32_t depth, v3* out) {
    v3 dir = reflect3(din, n);
    if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), fuzz));
    if (dot3(dir, n) <= 0.0f) return 0;
    *out = norm3(dir);
    return 1;
}
static int shade_metal(const or_scene* s, pstate* ps, const surf* f) { /* path_core.hpp:230-240 (+ fuzz) */
    const mtl* m = &s->M[f->material];
    v3 dir;
    if (!metal_dir(ps->r.d, f->n, m->fuzz, ps->key, (uint32_t)ps->depth, &dir)) return 0;
    ps->T = mulv(ps->T, m->albedo);
    ps->specular = 1;
    ps->prev_pdf = 0.0f;
    ps->depth += 1;
    next_ray(ps, add(f->p, muls(f->n, K_EPS)), dir);
    return roulette(ps);
}
int or_metal_dir(const float* d3, const float* n3, float fuzz, uint32_t k0, uint32_t k1, uint32_t depth, float* out3) {
    pkey k = {k0, k1};
    v3 o = mk(0.0f, 0.0f, 0.0f);
    int alive = metal_dir(mk(d3[0], d3[1], d3[2]), mk(n3[0], n3[1], n3[2]), fuzz, k, depth, &o);
    out3[0] = o.x; out3[1] = o.y; out3[2] = o.z;
    return alive;
}
static float fresnel(float ci, float ei, float et) { /* path_core.hpp:243-251 */
    float st = ei / et * sqrtf(fmax_r(0.0f, 1.0f - ci * ci));
    if (st >= 1.0f) return 1.0f;
    float ct = sqrtf(fmax_r(0.0f, 1.0f - st * st));
    float rpar = (et * ci - ei * ct) / (et * ci + ei * ct);
    float rperp = (ei * ci - et * ct) / (ei * ci + et * ct);
    return 0.5f * (rpar * rpar + rperp * rperp);
}
static int shade_dielectric(const or_scene* s, pstate* ps, const surf* f) { /* path_core.hpp:253-276 */
    const mtl* m = &s->M[f->material];
    float ei = f->inside ? m->ior : 1.0f, et = f->inside ? 1.0f : m->ior;
    float ci = clamp_r(-dot3(ps->r.d, f->n), 0.0f, 1.0f);
    float F = fresnel(ci, ei, et);
    float u = rnd(ps->key, (uint32_t)ps->depth, D_DIEL);
    v3 org, dir;
    if (u < F) {
        dir = reflect3(ps->r.d, f->n);
        org = add(f->p, muls(f->n, K_EPS));
    } else {
        float eta = ei / et;
        float s2 = eta * eta * (1.0f - ci * ci);
        float ct = sqrtf(fmax_r(0.0f, 1.0f - s2));
        dir = add(muls(ps->r.d, eta), muls(f->n, eta * ci - ct));
        org = sub(f->p, muls(f->n, K_EPS));
    }
    ps->T = mulv(ps->T, m->albedo);
    ps->specular = 1;
    ps->prev_pdf = 0.0f;
    ps->depth += 1;
    next_ray(ps, org, norm3(dir));
    return roulette(ps);
}
/* trace_path, path_core.hpp:279-304, with ray counting and the hit validity check of the GPU arms */
static v3 trace(const or_scene* s, int px, int py, uint32_t sample, uint32_t seed, int max_depth, uint64_t* cnt) {
    pstate ps = init_path(s, px, py, sample, seed);
    for (;;) {
        hitrec h;
        intersect(s, ps.r, &h, 0);
        cnt[0]++;
        if (!hit_valid(s, &h)) { cnt[2]++; h.prim = -1; }
        surf f;
        int mat = classify(s, &ps, &h, max_depth, &f);
        if (mat < 0) break;
        int alive;
        if (mat == LAMBERT) {
            ray sh; v3 c; int has;
            alive = shade_lambert(s, &ps, &f, &sh, &c, &has);
            if (has) {
                hitrec sh_h;
                cnt[1]++;
                if (!intersect(s, sh, &sh_h, 1)) ps.L = add(ps.L, c);
            }
        } else if (mat == METAL) {
            alive = shade_metal(s, &ps, &f);
        } else {
            alive = shade_dielectric(s, &ps, &f);
        }
        if (!alive) break;
    }
    return ps.L;
}

static void film_add(float* rgb, float* lum, float* lum2, v3 L) { /* src/gpu/engine.cuh:25-32 */
    rgb[0] += L.x; rgb[1] += L.y; rgb[2] += L.z;
    float l = lum3(L);
    *lum += l;
    *lum2 += l * l;
}

typedef struct {
    const or_scene* s;
    int first, count, max_depth;
    uint32_t seed;
    float *rgb, *lum, *lum2;
    int next_row;
    uint64_t cnt[3];
    pthread_mutex_t mu;
} rjob;

static void* render_worker(void* arg) {
    rjob* j = (rjob*)arg;
    uint64_t cnt[3] = {0, 0, 0};
    for (;;) {
        int py = __atomic_fetch_add(&j->next_row, 1, __ATOMIC_RELAXED);
        if (py >= j->s->H) break;
        for (int px = 0; px < j->s->W; ++px) {
            int p = py * j->s->W + px;
            for (int k = 0; k < j->count; ++k) {
                v3 L = trace(j->s, px, py, (uint32_t)(j->first + k), j->seed, j->max_depth, cnt);
                film_add(&j->rgb[3 * p], &j->lum[p], &j->lum2[p], L);
            }
        }
    }
    pthread_mutex_lock(&j->mu);
    for (int i = 0; i < 3; ++i) j->cnt[i] += cnt[i];
    pthread_mutex_unlock(&j->mu);
    return NULL;
}

void or_render(const or_scene* s, int first, int count, uint32_t seed, int max_depth, int threads, float* rgb,
               float* lum, float* lum2, uint64_t* counts) {
    rjob j;
    memset(&j, 0, sizeof(j));
    j.s = s; j.first = first; j.count = count; j.seed = seed; j.max_depth = max_depth;
    j.rgb = rgb; j.lum = lum; j.lum2 = lum2;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, render_worker, &j);
    render_worker(&j);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&j.mu);
    for (int i = 0; i < 3; ++i) counts[i] += j.cnt[i];
}

void or_trace_paths(const or_scene* s, const int* pxys, int n, uint32_t seed, int max_depth, float* L,
                    uint64_t* counts) {
    for (int i = 0; i < n; ++i) {
        v3 l = trace(s, pxys[3 * i], pxys[3 * i + 1], (uint32_t)pxys[3 * i + 2], seed, max_depth, counts);
        L[3 * i] = l.x; L[3 * i + 1] = l.y; L[3 * i + 2] = l.z;
    }
}

/* ---------------------------------------------------------------- wavefront pass (src/gpu/wavefront.cu:433-601) */
static uint64_t fnv_list(const uint32_t* v, int n) {
    uint64_t h = 14695981039346656037ULL;
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < 4; ++b) { h ^= (v[i] >> (8 * b)) & 0xffu; h *= 1099511628211ULL; }
    return h;
}

int or_wavefront_pass(const or_scene* s, uint32_t sample, uint32_t seed, int max_depth, int queue_capacity,
                      uint32_t* records, uint64_t* hashes, int max_records, float* rgb, float* lum, float* lum2,
                      uint64_t* counts) {
    const int P = s->W * s->H;
    const uint32_t cap = (queue_capacity > 0 && queue_capacity < P) ? (uint32_t)queue_capacity : (uint32_t)P;
    pstate* ps = (pstate*)malloc(sizeof(pstate) * (size_t)P);
    surf* sf = (surf*)malloc(sizeof(surf) * (size_t)P);
    hitrec* hit = (hitrec*)malloc(sizeof(hitrec) * (size_t)P);
    int* alive = (int*)malloc(sizeof(int) * (size_t)P);
    uint32_t* act = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)P);
    uint32_t* nxt = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)P);
    uint32_t* q[3];
    for (int k = 0; k < 3; ++k) q[k] = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)P);
    ray* shr = (ray*)malloc(sizeof(ray) * (size_t)P);
    v3* shc = (v3*)malloc(sizeof(v3) * (size_t)P);
    uint32_t* shp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)P);
    for (int i = 0; i < P; ++i) { ps[i] = init_path(s, i % s->W, i / s->W, sample, seed); act[i] = (uint32_t)i; } /* :48-70 */
    uint32_t n_act = (uint32_t)P;
    int bounce = 0;
    while (n_act > 0 && bounce < max_depth + 2) {
        uint32_t qn[4] = {0, 0, 0, 0};
        for (uint32_t k = 0; k < n_act; ++k) { /* extend :72-86 */
            uint32_t i = act[k];
            intersect(s, ps[i].r, &hit[i], 0);
        }
        counts[0] += n_act;
        for (uint32_t k = 0; k < n_act; ++k) { /* classify :88-130 */
            uint32_t i = act[k];
            if (!hit_valid(s, &hit[i])) { counts[2]++; hit[i].prim = -1; }
            int mat = classify(s, &ps[i], &hit[i], max_depth, &sf[i]);
            alive[i] = mat >= 0;
            if (mat >= 0) {
                uint32_t slot = qn[mat]++;
                if (slot < cap) q[mat][slot] = i;
                else alive[i] = 0; /* dropped on overflow: the path ends */
            }
        }
        for (int m = 0; m < 3; ++m) { /* shade :132-229 */
            uint32_t n = qn[m] < cap ? qn[m] : cap;
            for (uint32_t k = 0; k < n; ++k) {
                uint32_t i = q[m][k];
                if (m == LAMBERT) {
                    ray sh; v3 c; int has;
                    alive[i] = shade_lambert(s, &ps[i], &sf[i], &sh, &c, &has);
                    if (has) {
                        uint32_t slot = qn[3]++;
                        if (slot < cap) { shr[slot] = sh; shc[slot] = c; shp[slot] = i; }
                    }
                } else if (m == METAL) {
                    alive[i] = shade_metal(s, &ps[i], &sf[i]);
                } else {
                    alive[i] = shade_dielectric(s, &ps[i], &sf[i]);
                }
            }
        }
        uint32_t nsh = qn[3] < cap ? qn[3] : cap;
        for (uint32_t k = 0; k < nsh; ++k) { /* shadow :231-251 */
            hitrec h;
            if (!intersect(s, shr[k], &h, 1)) ps[shp[k]].L = add(ps[shp[k]].L, shc[k]);
        }
        counts[1] += nsh;
        uint32_t n2 = 0; /* mark + stable compaction :253-260, src/gpu/queues.cu:158-178 */
        for (uint32_t k = 0; k < n_act; ++k)
            if (alive[act[k]]) nxt[n2++] = act[k];
        if (bounce < max_records) {
            uint32_t* r = records + 6 * bounce;
            r[0] = n_act;
            for (int m = 0; m < 4; ++m) r[1 + m] = qn[m] < cap ? qn[m] : cap;
            r[5] = n2;
            hashes[bounce] = fnv_list(nxt, (int)n2);
        }
        uint32_t* t = act; act = nxt; nxt = t;
        n_act = n2;
        bounce++;
    }
    for (int i = 0; i < P; ++i) film_add(&rgb[3 * i], &lum[i], &lum2[i], ps[i].L); /* finish :263-270 */
    free(ps); free(sf); free(hit); free(alive); free(act); free(nxt);
    for (int k = 0; k < 3; ++k) free(q[k]);
    free(shr); free(shc); free(shp);
    return bounce;
}

void or_finalize(const float* rgb_sum, const float* lum_sum, const float* lum2_sum, int pixels, int spp, float* rgb,
                 float* lum_var) {
    for (int i = 0; i < pixels; ++i) {
        if (rgb)
            for (int c = 0; c < 3; ++c) rgb[3 * i + c] = rgb_sum[3 * i + c] / (float)spp;
        if (lum_var) {
            if (spp > 1) {
                double s1 = lum_sum[i], s2 = lum2_sum[i], n = spp;
                double v = (s2 - s1 * s1 / n) / (n - 1.0);
                lum_var[i] = (float)(v > 0.0 ? v : 0.0);
            } else {
                lum_var[i] = 0.0f;
            }
        }
    }
}

void or_compare(const float* a, const float* va, int spp_a, const float* b, const float* vb, int spp_b, int n,
                double* out) { /* src/film.cpp:45-113 */
    double max_rel = 0.0, ma = 0.0, mb = 0.0;
    int within = 0, nonfinite = 0;
    for (int i = 0; i < 3 * n; ++i)
        if (!isfinite(a[i]) || !isfinite(b[i])) nonfinite++;
    for (int i = 0; i < n; ++i) {
        float la = lum3(mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]));
        float lb = lum3(mk(b[3 * i], b[3 * i + 1], b[3 * i + 2]));
        ma += la;
        mb += lb;
        for (int c = 0; c < 3; ++c) {
            float x = a[3 * i + c], y = b[3 * i + c];
            if (isfinite(x) && isfinite(y)) {
                float m = fmax_r(fmax_r(fabsf(x), fabsf(y)), 1e-6f);
                double r = (double)(fabsf(x - y) / m);
                if (r > max_rel) max_rel = r;
            }
        }
        if (isfinite(la) && isfinite(lb)) {
            float sa = sqrtf(va[i] / (float)spp_a), sb = sqrtf(vb[i] / (float)spp_b);
            if (fabsf(la - lb) <= 3.0f * sqrtf(sa * sa + sb * sb)) within++;
        }
    }
    ma /= n;
    mb /= n;
    out[0] = max_rel;
    out[1] = ma > 1e-12 ? fabs(ma - mb) / ma : fabs(ma - mb);
    out[2] = n > 0 ? (double)within / n : 0.0;
    out[3] = nonfinite;
}

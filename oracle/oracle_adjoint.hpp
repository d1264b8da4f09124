// oracle_adjoint.hpp -- TEST INFRASTRUCTURE ONLY (the oracle): a CPU
// restatement of the path-replay adjoint, and an independent evaluator of the
// same per-path estimator with its discrete decisions frozen, for central
// finite differences.  Loaded by tests/ through liboracle.so; never by the
// product.
//
// The estimator (no reference counterpart: the reference renders a fixed scene,
// backends/metal/plugin.mm:584).  A path's radiance C is the integrator's sum
// (trace_path_nee / trace_path_reference in oracle.hpp) with every discrete
// decision held at its value for the base parameters: the triangle hit at each
// bounce, the facing side, the light chosen, the sign of the light-side cosine,
// the NEE sample's validity and visibility, the Russian-roulette outcome and
// weight 1/q, and the frame choice of the cosine-sampled direction.  The hit
// point of bounce j is the intersection of the incoming ray with the plane of
// the frozen triangle, x_j = o_j + t_j d_j, t_j = N.(P0 - o_j) / N.d_j; the next
// ray leaves from x_j + nf_j * 0.001 in the direction cosine_sample(nf_j) of the
// same two draws.  Light geometry (emitting triangles) is held fixed.
// The loss is l = sum_pixels dL/dI . (1/S) sum_s C_s.
//
// Parity pins: eval_frozen at the base parameters reproduces the renderer's
// image (tests/test_oracle_adjoint.py); adjoint_path agrees with central
// differences of eval_frozen; the albedo gradient satisfies Euler's identity
// for the degree of each NEE term, computed from forward renders alone.
#pragma once

#include "oracle.hpp"

namespace qo {

constexpr uint8_t kVxFacing = 1;     // dot(n, d) < 0: the facing normal is +n
constexpr uint8_t kVxNee = 2;        // the NEE sample was valid and unoccluded
constexpr uint8_t kVxSigmaNeg = 4;   // dot(n_light, w) < 0 for the NEE sample
constexpr uint8_t kVxFrameY = 8;     // the next direction's frame used a = (0, 1, 0)

// One hit vertex of a recorded path.
template <class R>
struct Seg {
    V3<R> o, d;      // incoming ray
    V3<R> thr;       // throughput before the vertex
    uint32_t tri;    // triangle hit (input order)
    uint32_t rng;    // RNG state before the vertex's own draws
    uint8_t flags;
    R qinv;          // Russian-roulette weight applied after the vertex (1 = none)
};

// cosine_sample with the frame choice given instead of tested.
template <class R>
inline V3<R> cosine_sample_frame(V3<R> normal, R r1, R r2, bool frame_y) {
    R phi = R(2) * pi<R>() * r1;
    R cos_t = std::sqrt(R(1) - r2);
    R sin_t = std::sqrt(r2);
    V3<R> w = normalize(normal);
    V3<R> a = frame_y ? V3<R>{0, 1, 0} : V3<R>{1, 0, 0};
    V3<R> v = normalize(cross(w, a));
    V3<R> u = cross(w, v);
    return normalize(u * (std::cos(phi) * sin_t) + v * (std::sin(phi) * sin_t) + w * cos_t);
}

// The path of the integrators from (o, d) with the RNG state after the camera
// draws, recorded vertex by vertex (same draws and decisions as
// trace_path_nee / trace_path_reference).
template <class R>
inline void record_path(const Scene& sc, const Bvh* bvh, V3<R> o, V3<R> d, uint32_t rng, const RenderParams& p,
                        std::vector<Seg<R>>& segs) {
    segs.clear();
    V3<R> thr{1, 1, 1};
    const bool nee = p.integrator == INTEGRATOR_NEE;
    const bool have_lights = !sc.light_tri.empty();
    for (uint32_t b = 0; b < p.max_depth; ++b) {
        Hit<R> h = closest_hit(sc, bvh, o, d, R(kTMin), R(kTFar), nullptr);
        if (h.tri == UINT32_MAX) break;
        Seg<R> sg{o, d, thr, h.tri, rng, 0, R(1)};
        V3<R> n = tri_normal<R>(sc, h.tri);
        R denom = dot(n, d);
        if (denom < R(0)) sg.flags |= kVxFacing;
        V3<R> nf = denom < R(0) ? n : -n;
        V3<R> pt = o + d * h.t;
        const float* e = &sc.emission[3 * size_t(h.tri)];
        const float* a = &sc.albedo[3 * size_t(h.tri)];
        V3<R> alb{R(a[0]), R(a[1]), R(a[2])};
        float emit_str = std::max(e[0], std::max(e[1], e[2]));
        if (emit_str > 0.1f || b + 1 >= p.max_depth) {
            segs.push_back(sg);
            break;
        }
        V3<R> origin = pt + nf * R(kOffset);
        if (nee && have_lights) {
            uint32_t li = sc.select_light(random_float(rng));
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            uint32_t lt = sc.light_tri[li];
            V3<R> l0 = sc.v0<R>(lt);
            V3<R> le1 = sc.v1<R>(lt) - l0;
            V3<R> le2 = sc.v2<R>(lt) - l0;
            R su = std::sqrt(r1);
            V3<R> xl = l0 + le1 * (su * (R(1) - r2)) + le2 * (su * r2);
            V3<R> w = xl - pt;
            R dist = std::sqrt(dot(w, w));
            V3<R> wi = w / dist;
            V3<R> nl = normalize(cross(le1, le2));
            R cos_x = dot(nf, wi);
            R cos_l = std::fabs(dot(nl, wi));
            if (dot(nl, wi) < R(0)) sg.flags |= kVxSigmaNeg;
            if (cos_x > R(0) && cos_l > R(0)) {
                V3<R> sv = xl - origin;
                R sl = std::sqrt(dot(sv, sv));
                if (!any_hit(sc, bvh, origin, sv / sl, R(kTMin), sl * R(1.0 - 0.001), nullptr)) sg.flags |= kVxNee;
            }
        }
        thr = thr * alb;
        if (b >= p.rr_start) {
            R q = std::max(R(0.05), max3(thr));
            if (R(random_float(rng)) > q) {
                segs.push_back(sg);
                break;
            }
            thr = thr / q;
            sg.qinv = R(1) / q;
        }
        if (std::fabs(normalize(nf).x) > R(0.9)) sg.flags |= kVxFrameY;
        segs.push_back(sg);
        o = origin;
        d = cosine_sample(nf, rng);
    }
}

// Triangle corners and albedos of an evaluation (base or perturbed).
template <class R>
struct ParamView {
    const Scene* sc;
    const double* corners;   // 9 per triangle (P0, P1, P2), or null for the scene's own
    const double* albedo;    // 3 per triangle, or null for the scene's own
    V3<R> p(uint32_t t, int k) const {
        if (corners) {
            const double* c = corners + 9 * size_t(t) + 3 * k;
            return {R(c[0]), R(c[1]), R(c[2])};
        }
        return k == 0 ? sc->v0<R>(t) : (k == 1 ? sc->v1<R>(t) : sc->v2<R>(t));
    }
    V3<R> alb(uint32_t t) const {
        if (albedo) return {R(albedo[3 * t]), R(albedo[3 * t + 1]), R(albedo[3 * t + 2])};
        const float* a = &sc->albedo[3 * size_t(t)];
        return {R(a[0]), R(a[1]), R(a[2])};
    }
};

// Radiance of a path whose decisions are `tape` (tri, flags, qinv of each
// vertex) for the parameters of `pv`; (o, d, rng) is the camera ray and the
// RNG state after the camera draws.
template <class R, class Vx>
inline V3<R> eval_frozen(const ParamView<R>& pv, const std::vector<Vx>& tape, V3<R> o, V3<R> d, uint32_t rng,
                         const RenderParams& p) {
    const Scene& sc = *pv.sc;
    const bool nee = p.integrator == INTEGRATOR_NEE;
    const bool have_lights = !sc.light_tri.empty();
    V3<R> thr{1, 1, 1}, C{0, 0, 0};
    for (size_t j = 0; j < tape.size(); ++j) {
        const uint32_t b = uint32_t(j);
        const uint32_t t = tape[j].tri;
        const uint8_t fl = tape[j].flags;
        V3<R> P0 = pv.p(t, 0), P1 = pv.p(t, 1), P2 = pv.p(t, 2);
        V3<R> N = cross(P1 - P0, P2 - P0);
        V3<R> n = normalize(N);
        V3<R> nf = (fl & kVxFacing) ? n : -n;
        R th = dot(N, P0 - o) / dot(N, d);
        V3<R> x = o + d * th;
        const float* e = &sc.emission[3 * size_t(t)];
        V3<R> Le{R(e[0]), R(e[1]), R(e[2])};
        V3<R> alb = pv.alb(t);
        if (!nee || b == 0) C = C + thr * Le;
        if (std::max(e[0], std::max(e[1], e[2])) > 0.1f || b + 1 >= p.max_depth) break;
        V3<R> origin = x + nf * R(kOffset);
        if (nee && have_lights) {
            uint32_t li = sc.select_light(random_float(rng));
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            if (fl & kVxNee) {
                uint32_t lt = sc.light_tri[li];
                V3<R> l0 = sc.v0<R>(lt);
                V3<R> le1 = sc.v1<R>(lt) - l0;
                V3<R> le2 = sc.v2<R>(lt) - l0;
                R su = std::sqrt(r1);
                V3<R> xl = l0 + le1 * (su * (R(1) - r2)) + le2 * (su * r2);
                V3<R> w = xl - x;
                R dist2 = dot(w, w);
                R dist = std::sqrt(dist2);
                V3<R> nl = normalize(cross(le1, le2));
                R cos_x = dot(nf, w) / dist;
                R cos_l = ((fl & kVxSigmaNeg) ? -dot(nl, w) : dot(nl, w)) / dist;
                const float* el = &sc.emission[3 * size_t(lt)];
                V3<R> Ll{R(el[0]), R(el[1]), R(el[2])};
                R g = cos_x * cos_l * R(sc.light_area[li]) / (pi<R>() * dist2 * R(sc.light_psel[li]));
                C = C + thr * alb * Ll * g;
            }
        }
        thr = thr * alb;
        if (b >= p.rr_start) {
            (void)random_float(rng);
            thr = thr * R(tape[j].qinv);
        }
        if (j + 1 < tape.size()) {
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            d = cosine_sample_frame(nf, r1, r2, (fl & kVxFrameY) != 0);
            o = origin;
        }
    }
    return C;
}

// Vector-Jacobian product of cosine_sample_frame's direction with respect to
// its normal.
template <class R>
inline V3<R> cosine_sample_vjp(V3<R> n, R r1, R r2, bool frame_y, V3<R> gbar) {
    R phi = R(2) * pi<R>() * r1;
    R cos_t = std::sqrt(R(1) - r2);
    R sin_t = std::sqrt(r2);
    R cp = std::cos(phi), sp = std::sin(phi);
    R ln = std::sqrt(dot(n, n));
    V3<R> w = n / ln;
    V3<R> a = frame_y ? V3<R>{0, 1, 0} : V3<R>{1, 0, 0};
    V3<R> vv = cross(w, a);
    R lv = std::sqrt(dot(vv, vv));
    V3<R> v = vv / lv;
    V3<R> u = cross(w, v);
    V3<R> D = u * (cp * sin_t) + v * (sp * sin_t) + w * cos_t;
    R lD = std::sqrt(dot(D, D));
    V3<R> dir = D / lD;
    V3<R> Db = (gbar - dir * dot(dir, gbar)) / lD;
    V3<R> ub = Db * (cp * sin_t), vb = Db * (sp * sin_t), wb = Db * cos_t;
    wb = wb + cross(v, ub);
    vb = vb + cross(ub, w);
    V3<R> Vb = (vb - v * dot(v, vb)) / lv;
    wb = wb + cross(a, Vb);
    return (wb - w * dot(w, wb)) / ln;
}

// Reverse mode over one recorded path for the loss weight delta (dL/dI / S):
// adds dl/dalbedo (3 per triangle) and dl/dcorner (9 per triangle).
template <class R>
inline void adjoint_path(const Scene& sc, const std::vector<Seg<R>>& segs, const RenderParams& p, V3<R> delta,
                         double* g_alb, double* g_corner) {
    const bool nee = p.integrator == INTEGRATOR_NEE;
    const bool have_lights = !sc.light_tri.empty();
    const int K = int(segs.size());
    V3<R> V{0, 0, 0}, ob{0, 0, 0}, db{0, 0, 0};
    for (int j = K - 1; j >= 0; --j) {
        const Seg<R>& sg = segs[size_t(j)];
        const uint32_t b = uint32_t(j);
        const uint32_t t = sg.tri;
        const float* e = &sc.emission[3 * size_t(t)];
        const float* a = &sc.albedo[3 * size_t(t)];
        V3<R> Le{R(e[0]), R(e[1]), R(e[2])};
        V3<R> alb{R(a[0]), R(a[1]), R(a[2])};
        V3<R> E = (!nee || b == 0) ? Le : V3<R>{0, 0, 0};
        if (std::max(e[0], std::max(e[1], e[2])) > 0.1f) {
            V = E;
            ob = db = V3<R>{0, 0, 0};
            continue;
        }
        V3<R> P0 = sc.v0<R>(t), e1 = sc.v1<R>(t) - P0, e2 = sc.v2<R>(t) - P0;
        V3<R> N = cross(e1, e2);
        R lN = std::sqrt(dot(N, N));
        V3<R> n = N / lN;
        R sgn = (sg.flags & kVxFacing) ? R(1) : R(-1);
        V3<R> nf = n * sgn;
        R den = dot(N, sg.d);
        R th = dot(N, P0 - sg.o) / den;
        V3<R> x = sg.o + sg.d * th;
        V3<R> xb = ob;
        V3<R> nfb = ob * R(kOffset);
        V3<R> Nt{0, 0, 0};
        uint32_t rng = sg.rng;
        const bool more = b + 1 < p.max_depth;
        if (nee && more && have_lights) {
            uint32_t li = sc.select_light(random_float(rng));
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            if (sg.flags & kVxNee) {
                uint32_t lt = sc.light_tri[li];
                V3<R> l0 = sc.v0<R>(lt);
                V3<R> le1 = sc.v1<R>(lt) - l0;
                V3<R> le2 = sc.v2<R>(lt) - l0;
                R su = std::sqrt(r1);
                V3<R> xl = l0 + le1 * (su * (R(1) - r2)) + le2 * (su * r2);
                V3<R> nl = normalize(cross(le1, le2));
                const float* el = &sc.emission[3 * size_t(lt)];
                V3<R> Ll{R(el[0]), R(el[1]), R(el[2])};
                V3<R> w = xl - x;
                R w2 = dot(w, w);
                R av = dot(nf, w);
                R sig = (sg.flags & kVxSigmaNeg) ? R(-1) : R(1);
                R bb = sig * dot(nl, w);
                R c = R(sc.light_area[li]) / (pi<R>() * R(sc.light_psel[li]));
                R i4 = R(1) / (w2 * w2);
                R gg = c * av * bb * i4;
                Nt = Ll * gg;
                V3<R> wt = delta * sg.thr * alb;
                R gb = dot(wt, Ll);
                nfb = nfb + w * (gb * c * bb * i4);
                V3<R> dgdw = (nf * bb + nl * (sig * av)) * (c * i4) - w * (R(4) * gg / w2);
                xb = xb - dgdw * gb;
            }
        }
        R qinv = R(1);
        if (more && b >= p.rr_start) {
            qinv = sg.qinv;
            (void)random_float(rng);
        }
        if (j + 1 < K) {
            R r1 = R(random_float(rng));
            R r2 = R(random_float(rng));
            nfb = nfb + cosine_sample_vjp(nf, r1, r2, (sg.flags & kVxFrameY) != 0, db);
        }
        V3<R> U = Nt + V * qinv;
        V3<R> gal = delta * sg.thr * U;
        V = E + alb * U;
        V3<R> nb = nfb * sgn;
        V3<R> Nb = (nb - n * dot(n, nb)) / lN;
        R tb = dot(xb, sg.d);
        V3<R> obn = xb;
        V3<R> dbn = xb * th;
        R numb = tb / den;
        R denb = -tb * th / den;
        Nb = Nb + (P0 - sg.o) * numb + sg.d * denb;
        V3<R> P0b = N * numb;
        obn = obn - N * numb;
        dbn = dbn + N * denb;
        V3<R> e1b = cross(e2, Nb);
        V3<R> e2b = cross(Nb, e1);
        P0b = P0b - e1b - e2b;
        ob = obn;
        db = dbn;
        double* ga = g_alb + 3 * size_t(t);
        ga[0] += double(gal.x); ga[1] += double(gal.y); ga[2] += double(gal.z);
        double* gc = g_corner + 9 * size_t(t);
        const V3<R> cs[3] = {P0b, e1b, e2b};
        for (int k = 0; k < 3; ++k) {
            gc[3 * k] += double(cs[k].x);
            gc[3 * k + 1] += double(cs[k].y);
            gc[3 * k + 2] += double(cs[k].z);
        }
    }
}

// Camera ray and RNG state (after the jitter draws) of pixel (x, y), sample s.
template <class R>
inline void camera_path_start(const Camera<R>& cam, const RenderParams& p, uint32_t x, uint32_t y, uint32_t s,
                              V3<R>& o, V3<R>& d, uint32_t& rng) {
    rng = p.rng_mode == RNG_REFERENCE ? seed_reference(x, y, s) : seed_hashed(x, y, p.width, p.height, s, p.seed);
    R u, v;
    pixel_uv<R>(x, y, p.width, p.height, rng, u, v);
    o = cam.origin;
    d = cam.direction(u, v);
}

// dl/dalbedo (3 per triangle) and dl/dcorner (9 per triangle) over the image:
// pixels whose dimg (RGBA per pixel) is zero, or whose index is not a multiple
// of p.pixel_stride, are skipped.  Multithreaded over rows.
template <class R>
inline void adjoint_render(const Scene& sc, const Bvh* bvh, const RenderParams& p, const float* dimg,
                           std::vector<double>& g_alb, std::vector<double>& g_corner) {
    Camera<R> cam = Camera<R>::make(p.camera, p.width, p.height);
    const uint32_t W = p.width, H = p.height;
    const uint32_t stride = std::max(1u, p.pixel_stride);
    const R scale = R(1) / R(p.spp);
    unsigned nthreads = p.threads ? p.threads : std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<double>> ta(nthreads, std::vector<double>(3 * size_t(sc.ntri), 0.0));
    std::vector<std::vector<double>> tc(nthreads, std::vector<double>(9 * size_t(sc.ntri), 0.0));
    std::atomic<uint32_t> next_row{0};
    auto worker = [&](unsigned tid) {
        std::vector<Seg<R>> segs;
        for (;;) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= H) break;
            for (uint32_t x = 0; x < W; ++x) {
                size_t pix = size_t(y) * W + x;
                if (pix % stride) continue;
                V3<R> delta{R(dimg[4 * pix]) * scale, R(dimg[4 * pix + 1]) * scale, R(dimg[4 * pix + 2]) * scale};
                if (delta.x == R(0) && delta.y == R(0) && delta.z == R(0)) continue;
                for (uint32_t s = p.sample_offset; s < p.sample_offset + p.spp; ++s) {
                    V3<R> o, d;
                    uint32_t rng;
                    camera_path_start(cam, p, x, y, s, o, d, rng);
                    record_path(sc, bvh, o, d, rng, p, segs);
                    adjoint_path(sc, segs, p, delta, ta[tid].data(), tc[tid].data());
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    g_alb.assign(3 * size_t(sc.ntri), 0.0);
    g_corner.assign(9 * size_t(sc.ntri), 0.0);
    for (unsigned t = 0; t < nthreads; ++t) {
        for (size_t i = 0; i < g_alb.size(); ++i) g_alb[i] += ta[t][i];
        for (size_t i = 0; i < g_corner.size(); ++i) g_corner[i] += tc[t][i];
    }
}

// The decisions of every path of an image, recorded once in FP64, and the
// loss weight of its pixel, for repeated frozen evaluations.
struct TapeVx {
    uint32_t tri;
    uint8_t flags;
    double qinv;
};

struct Tape {
    RenderParams p;
    std::vector<uint32_t> path_pixel, path_sample, path_first, path_count;
    std::vector<double> path_delta;   // 3 per path: dL/dI / S
    std::vector<TapeVx> vx;

    void record(const Scene& sc, const Bvh* bvh, const RenderParams& rp, const float* dimg) {
        p = rp;
        Camera<double> cam = Camera<double>::make(p.camera, p.width, p.height);
        const uint32_t stride = std::max(1u, p.pixel_stride);
        std::vector<Seg<double>> segs;
        for (uint32_t y = 0; y < p.height; ++y)
            for (uint32_t x = 0; x < p.width; ++x) {
                size_t pix = size_t(y) * p.width + x;
                if (pix % stride) continue;
                double dl[3] = {double(dimg[4 * pix]) / p.spp, double(dimg[4 * pix + 1]) / p.spp,
                                double(dimg[4 * pix + 2]) / p.spp};
                if (dl[0] == 0.0 && dl[1] == 0.0 && dl[2] == 0.0) continue;
                for (uint32_t s = p.sample_offset; s < p.sample_offset + p.spp; ++s) {
                    V3<double> o, d;
                    uint32_t rng;
                    camera_path_start(cam, p, x, y, s, o, d, rng);
                    record_path(sc, bvh, o, d, rng, p, segs);
                    path_pixel.push_back(uint32_t(pix));
                    path_sample.push_back(s);
                    path_first.push_back(uint32_t(vx.size()));
                    path_count.push_back(uint32_t(segs.size()));
                    for (int k = 0; k < 3; ++k) path_delta.push_back(dl[k]);
                    for (const Seg<double>& sg : segs) vx.push_back(TapeVx{sg.tri, sg.flags, sg.qinv});
                }
            }
    }

    // l = sum over the paths that visit a triangle with relevant[t] != 0 (all
    // paths when relevant is null) of delta . C for the parameters of pv.
    double eval(const ParamView<double>& pv, const std::vector<char>* relevant, unsigned threads) const {
        Camera<double> cam = Camera<double>::make(p.camera, p.width, p.height);
        const size_t np = path_pixel.size();
        const size_t chunk = 4096;
        const size_t nchunks = (np + chunk - 1) / chunk;
        unsigned nth = std::max(1u, threads);
        // one partial per chunk, summed in chunk order: the result does not
        // depend on the thread count or scheduling
        std::vector<double> part(nchunks, 0.0);
        std::atomic<size_t> next{0};
        auto worker = [&](unsigned) {
            std::vector<TapeVx> tape;
            for (;;) {
                size_t c = next.fetch_add(1);
                if (c >= nchunks) break;
                size_t b0 = c * chunk, b1 = std::min(np, b0 + chunk);
                for (size_t i = b0; i < b1; ++i) {
                    const TapeVx* v0 = vx.data() + path_first[i];
                    const uint32_t cnt = path_count[i];
                    if (relevant) {
                        bool hit = false;
                        for (uint32_t k = 0; k < cnt && !hit; ++k) hit = (*relevant)[v0[k].tri] != 0;
                        if (!hit) continue;
                    }
                    tape.assign(v0, v0 + cnt);
                    uint32_t pix = path_pixel[i];
                    V3<double> o, d;
                    uint32_t rng;
                    camera_path_start(cam, p, pix % p.width, pix / p.width, path_sample[i], o, d, rng);
                    V3<double> C = eval_frozen(pv, tape, o, d, rng, p);
                    part[c] += path_delta[3 * i] * C.x + path_delta[3 * i + 1] * C.y + path_delta[3 * i + 2] * C.z;
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nth; ++t) pool.emplace_back(worker, t);
        worker(0);
        for (auto& th : pool) th.join();
        double l = 0.0;
        for (double v : part) l += v;
        return l;
    }
};

}  // namespace qo

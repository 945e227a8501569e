// Host-side workload synthesis for the visibility benchmarks: rasterises
// axis-aligned occluder boxes into an f32 depth buffer by casting one ray per
// pixel centre (nearest hit), in double precision, then storing the NDC depth
// of the hit point as f32. Used to produce the Hi-Z input depth of the
// occlusion workloads; not part of the culling path.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

struct V3 { double x, y, z; };

inline void mul4(const double* m, double x, double y, double z, double w, double out[4]) {
    for (int r = 0; r < 4; ++r) out[r] = m[r] * x + m[4 + r] * y + m[8 + r] * z + m[12 + r] * w;
}

}  // namespace

extern "C" {

// view_proj / inv_view_proj: column-major doubles. boxes: n x {min xyz, max xyz}
// as f32. depth: w x h f32 row-major (row 0 at the top), pre-filled by the
// caller with the clear value; hits overwrite it when nearer. reversed: 1 when
// NDC z = 1 is the near plane.
int hvs_rasterize_boxes(const double* view_proj, const double* inv_view_proj, const double* cam, uint32_t w,
                        uint32_t h, const float* boxes, uint32_t n, int reversed, float* depth, uint32_t threads) {
    if (!view_proj || !inv_view_proj || !cam || !boxes || !depth || !w || !h) return -1;
    // Screen-space bounding rectangles of every box (whole screen if a corner is behind the eye).
    struct Rect { int x0, y0, x1, y1; };
    std::vector<Rect> rects(n);
    for (uint32_t b = 0; b < n; ++b) {
        const float* bx = boxes + 6 * (size_t)b;
        // A box around the eye is seen from inside (back faces only): it
        // occludes nothing, as with back-face culling in a depth prepass.
        if (cam[0] >= bx[0] && cam[0] <= bx[3] && cam[1] >= bx[1] && cam[1] <= bx[4] && cam[2] >= bx[2] &&
            cam[2] <= bx[5]) {
            rects[b] = {1, 1, 0, 0};
            continue;
        }
        double mnx = 1e30, mny = 1e30, mxx = -1e30, mxy = -1e30;
        bool behind = false;
        for (int c = 0; c < 8; ++c) {
            double p[4];
            mul4(view_proj, (c & 1) ? bx[3] : bx[0], (c & 2) ? bx[4] : bx[1], (c & 4) ? bx[5] : bx[2], 1.0, p);
            if (p[3] <= 1e-9) { behind = true; break; }
            const double sx = (p[0] / p[3] * 0.5 + 0.5) * w, sy = (0.5 - p[1] / p[3] * 0.5) * h;
            mnx = std::min(mnx, sx); mxx = std::max(mxx, sx);
            mny = std::min(mny, sy); mxy = std::max(mxy, sy);
        }
        Rect r;
        if (behind) {
            r = {0, 0, (int)w - 1, (int)h - 1};
        } else {
            r.x0 = (int)std::max(0.0, std::floor(mnx) - 1);
            r.y0 = (int)std::max(0.0, std::floor(mny) - 1);
            r.x1 = (int)std::min((double)w - 1, std::ceil(mxx) + 1);
            r.y1 = (int)std::min((double)h - 1, std::ceil(mxy) + 1);
        }
        rects[b] = r;
    }
    std::vector<double> tbest((size_t)w * h, 1e300);
    auto work = [&](uint32_t row0, uint32_t row1) {
        for (uint32_t b = 0; b < n; ++b) {
            const Rect& r = rects[b];
            if (r.x1 < r.x0 || r.y1 < r.y0) continue;
            const float* bx = boxes + 6 * (size_t)b;
            const int ya = std::max<int>(r.y0, (int)row0), yb = std::min<int>(r.y1, (int)row1 - 1);
            for (int y = ya; y <= yb; ++y) {
                const double ny = 1.0 - (y + 0.5) / h * 2.0;
                for (int x = r.x0; x <= r.x1; ++x) {
                    const double nx = (x + 0.5) / w * 2.0 - 1.0;
                    double a[4], c[4];
                    mul4(inv_view_proj, nx, ny, 0.0, 1.0, a);
                    mul4(inv_view_proj, nx, ny, 1.0, 1.0, c);
                    const V3 pa{a[0] / a[3], a[1] / a[3], a[2] / a[3]};
                    const V3 pc{c[0] / c[3], c[1] / c[3], c[2] / c[3]};
                    V3 d{pc.x - pa.x, pc.y - pa.y, pc.z - pa.z};
                    const double len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
                    d = {d.x / len, d.y / len, d.z / len};
                    // Ray from the eye; keep the direction pointing away from it.
                    const V3 o{cam[0], cam[1], cam[2]};
                    const double dot = (pa.x - o.x) * d.x + (pa.y - o.y) * d.y + (pa.z - o.z) * d.z +
                                       (pc.x - o.x) * d.x + (pc.y - o.y) * d.y + (pc.z - o.z) * d.z;
                    if (dot < 0) d = {-d.x, -d.y, -d.z};
                    double t0 = 0.0, t1 = 1e300;
                    const double od[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
                    bool hit = true;
                    for (int k = 0; k < 3 && hit; ++k) {
                        if (std::fabs(dd[k]) < 1e-300) {
                            if (od[k] < bx[k] || od[k] > bx[3 + k]) hit = false;
                        } else {
                            double ta = (bx[k] - od[k]) / dd[k], tb = (bx[3 + k] - od[k]) / dd[k];
                            if (ta > tb) std::swap(ta, tb);
                            t0 = std::max(t0, ta);
                            t1 = std::min(t1, tb);
                            if (t0 > t1) hit = false;
                        }
                    }
                    if (!hit) continue;
                    const size_t idx = (size_t)y * w + x;
                    if (t0 < tbest[idx]) tbest[idx] = t0;
                }
            }
        }
    };
    if (threads < 1) threads = 1;
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) {
        const uint32_t r0 = (uint32_t)((uint64_t)h * t / threads), r1 = (uint32_t)((uint64_t)h * (t + 1) / threads);
        pool.emplace_back(work, r0, r1);
    }
    for (auto& th : pool) th.join();
    for (uint32_t y = 0; y < h; ++y) {
        const double ny = 1.0 - (y + 0.5) / h * 2.0;
        for (uint32_t x = 0; x < w; ++x) {
            const size_t idx = (size_t)y * w + x;
            if (tbest[idx] >= 1e299) continue;
            const double nx = (x + 0.5) / w * 2.0 - 1.0;
            double a[4], c[4];
            mul4(inv_view_proj, nx, ny, 0.0, 1.0, a);
            mul4(inv_view_proj, nx, ny, 1.0, 1.0, c);
            V3 pa{a[0] / a[3], a[1] / a[3], a[2] / a[3]}, pc{c[0] / c[3], c[1] / c[3], c[2] / c[3]};
            V3 d{pc.x - pa.x, pc.y - pa.y, pc.z - pa.z};
            const double len = std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
            d = {d.x / len, d.y / len, d.z / len};
            const double dot = (pa.x - cam[0]) * d.x + (pa.y - cam[1]) * d.y + (pa.z - cam[2]) * d.z +
                               (pc.x - cam[0]) * d.x + (pc.y - cam[1]) * d.y + (pc.z - cam[2]) * d.z;
            if (dot < 0) d = {-d.x, -d.y, -d.z};
            const double t = tbest[idx];
            double p[4];
            mul4(view_proj, cam[0] + d.x * t, cam[1] + d.y * t, cam[2] + d.z * t, 1.0, p);
            if (p[3] <= 0) continue;
            double z = p[2] / p[3];
            z = std::min(1.0, std::max(0.0, z));
            const float zf = (float)z;
            const bool nearer = reversed ? (zf > depth[idx]) : (zf < depth[idx]);
            if (nearer) depth[idx] = zf;
        }
    }
    return 0;
}

}  // extern "C"

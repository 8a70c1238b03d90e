/*
 * tiled_cpu.c — C/OpenMP restatement of one CT-splatting training iteration
 * (projection -> tile binning -> rasterize -> MSE adjoint -> rasterize VJP ->
 * projection/parameter VJPs) in float32. TEST INFRASTRUCTURE ONLY: it is the
 * timed CPU baseline of bench.py and is checked against the float64 oracle in
 * tests; the product never calls it.
 *
 * Restates (reference repository paths):
 *   parameter transform   r2splat/cuda/csrc/ParameterTransformFwd.cu:10-40, ParameterTransformBwd.cu:10-77
 *   projection            r2splat/cuda/csrc/WorldToImageProjectionFwd.cu:73-283, ...Bwd.cu:67-554
 *   conics / mu           r2splat/cuda/csrc/ConicsMuFwd.cu:35-104, ConicsMuBwd.cu:44-129
 *   tile bounds           r2splat/cuda/csrc/IsectTilesKernel.cu:30-112 (stable per-tile lists)
 *   rasterize fwd / VJP   r2splat/cuda/csrc/RasterizeToPixelsFwd.cu:155-180, RasterizeToPixelsBwd.cu:421-469
 * Parallel over cameras (one camera per OpenMP thread at a time); per-camera
 * gradient partials are summed in camera order, so results are deterministic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    float mx, my, a, b, c, opac, mu;
    int valid;
    /* saved for the VJP */
    float R[9], fx, fy, x, y, z, tx, ty, l, J[9], Sc[6], c2[6];
    int clip_x, clip_y;
} Proj;

static void quat_rot(const float* raw, float q[4], float* mag, float* norm, float R[9]) {
    float m = 0.f;
    for (int i = 0; i < 4; ++i) m = fmaxf(m, fabsf(raw[i]));
    float ss = 0.f;
    for (int i = 0; i < 4; ++i) {
        q[i] = raw[i] / m;
        ss += q[i] * q[i];
    }
    float n = sqrtf(ss);
    for (int i = 0; i < 4; ++i) q[i] /= n;
    *mag = m;
    *norm = n;
    float w = q[0], x = q[1], y = q[2], z = q[3];
    R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
    R[3] = 2 * (x * y + w * z); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - w * x);
    R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = 1 - 2 * (x * x + y * y);
}

static void mat3_mul(const float* A, const float* B, float* C) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

static void mat3_t(const float* A, float* T) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) T[3 * j + i] = A[3 * i + j];
}

static void sym_full(const float* p, float* M) {
    M[0] = p[0]; M[1] = p[1]; M[2] = p[2];
    M[3] = p[1]; M[4] = p[3]; M[5] = p[4];
    M[6] = p[2]; M[7] = p[4]; M[8] = p[5];
}

static void project_one(const float* view, const float* K, const float* p, const float* cov, int W, int H,
                        float eps, float clip, Proj* o) {
    memset(o, 0, sizeof(*o));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o->R[3 * i + j] = view[4 * i + j];
    float x = o->R[0] * p[0] + o->R[1] * p[1] + o->R[2] * p[2] + view[3];
    float y = o->R[3] * p[0] + o->R[4] * p[1] + o->R[5] * p[2] + view[7];
    float z = o->R[6] * p[0] + o->R[7] * p[1] + o->R[8] * p[2] + view[11];
    o->x = x; o->y = y; o->z = z;
    o->fx = K[0]; o->fy = K[4];
    if (!(z > 0.01f)) return;
    float cx = K[2], cy = K[5];
    float lxp = (1.15f * W - cx) / o->fx, lxn = (cx + 0.15f * W) / o->fx;
    float lyp = (1.15f * H - cy) / o->fy, lyn = (cy + 0.15f * H) / o->fy;
    float u = x / z, v = y / z;
    o->clip_x = (u > lxp) || (u < -lxn);
    o->clip_y = (v > lyp) || (v < -lyn);
    o->tx = z * fminf(lxp, fmaxf(-lxn, u));
    o->ty = z * fminf(lyp, fmaxf(-lyn, v));
    o->l = sqrtf(o->tx * o->tx + o->ty * o->ty + z * z);
    float J[9] = {o->fx / z, 0, -o->fx * o->tx / (z * z), 0, o->fy / z, -o->fy * o->ty / (z * z),
                  o->tx / o->l, o->ty / o->l, z / o->l};
    memcpy(o->J, J, sizeof(J));
    float S[9], RS[9], Rt[9], Scf[9], JS[9], Jt[9], C2[9];
    sym_full(cov, S);
    mat3_mul(o->R, S, RS);
    mat3_t(o->R, Rt);
    mat3_mul(RS, Rt, Scf);
    o->Sc[0] = Scf[0]; o->Sc[1] = Scf[1]; o->Sc[2] = Scf[2]; o->Sc[3] = Scf[4]; o->Sc[4] = Scf[5]; o->Sc[5] = Scf[8];
    mat3_mul(J, Scf, JS);
    mat3_t(J, Jt);
    mat3_mul(JS, Jt, C2);
    o->c2[0] = C2[0]; o->c2[1] = C2[1]; o->c2[2] = C2[2]; o->c2[3] = C2[4]; o->c2[4] = C2[5]; o->c2[5] = C2[8];
    o->mx = o->fx * x / z + cx;
    o->my = o->fy * y / z + cy;
    /* conic and path factor */
    double a = (double)o->c2[0] + eps, b = o->c2[1], d = (double)o->c2[3] + eps;
    double det = a * d - b * b;
    if (!(a > 0 && d > 0 && det > 0)) return;
    double P0 = d / det, P1 = -b / det, P2 = a / det;
    double w0 = P0 * o->c2[2] + P1 * o->c2[4], w1 = P1 * o->c2[2] + P2 * o->c2[4];
    double s = (double)o->c2[5] + eps - o->c2[2] * w0 - o->c2[4] * w1;
    o->a = (float)P0; o->b = (float)P1; o->c = (float)P2;
    o->mu = s > 0 ? (float)sqrt(2.0 * M_PI * s) : 0.f;
    (void)clip;
    o->valid = 1;
}

static int span(float mean, int r, int tile, int n, int* lo, int* hi) {
    float c = mean / (float)tile, rr = (float)r / (float)tile;
    float l = floorf(c - rr), h = ceilf(c + rr);
    *lo = (!isfinite(l) || l <= 0) ? 0 : (l >= n ? n : (int)l);
    *hi = (!isfinite(h) || h <= 0) ? 0 : (h >= n ? n : (int)h);
    return *hi > *lo;
}

/* Projection VJP for one pair (same chain rule as the GPU kernel). */
static void proj_vjp(const Proj* o, const float vm[2], const float vc[6], float vp[3], float vcov[6]) {
    const float* J = o->J;
    float G[9] = {vc[0], vc[1], vc[2], 0, vc[3], vc[4], 0, 0, vc[5]};
    float GJ[9], Jt[9], vSc[9], T1[9], Rt[9], vS[9];
    mat3_mul(G, J, GJ);
    mat3_t(J, Jt);
    mat3_mul(Jt, GJ, vSc);
    mat3_mul(vSc, o->R, T1);
    mat3_t(o->R, Rt);
    mat3_mul(Rt, T1, vS);
    vcov[0] += vS[0]; vcov[1] += vS[1] + vS[3]; vcov[2] += vS[2] + vS[6];
    vcov[3] += vS[4]; vcov[4] += vS[5] + vS[7]; vcov[5] += vS[8];
    float Sc[9];
    sym_full(o->Sc, Sc);
    float JSc[9], GG[9] = {2 * vc[0], vc[1], vc[2], vc[1], 2 * vc[3], vc[4], vc[2], vc[4], 2 * vc[5]}, vJ[9];
    mat3_mul(J, Sc, JSc);
    mat3_mul(GG, JSc, vJ);
    float z = o->z, rz = 1.f / z, rz2 = rz * rz, rz3 = rz2 * rz, rl = 1.f / o->l, rl3 = rl * rl * rl;
    float tx = o->tx, ty = o->ty;
    float v_tx = vJ[2] * (-o->fx * rz2) + vJ[6] * (rl - tx * tx * rl3) + vJ[7] * (-tx * ty * rl3) + vJ[8] * (-z * tx * rl3);
    float v_ty = vJ[5] * (-o->fy * rz2) + vJ[6] * (-tx * ty * rl3) + vJ[7] * (rl - ty * ty * rl3) + vJ[8] * (-z * ty * rl3);
    float v_z = vJ[0] * (-o->fx * rz2) + vJ[2] * (2 * o->fx * tx * rz3) + vJ[4] * (-o->fy * rz2) +
                vJ[5] * (2 * o->fy * ty * rz3) + vJ[6] * (-tx * z * rl3) + vJ[7] * (-ty * z * rl3) + vJ[8] * (rl - z * z * rl3);
    float v_x = 0, v_y = 0;
    if (o->clip_x) v_z += v_tx * tx * rz; else v_x += v_tx;
    if (o->clip_y) v_z += v_ty * ty * rz; else v_y += v_ty;
    v_x += o->fx * rz * vm[0];
    v_y += o->fy * rz * vm[1];
    v_z -= (o->fx * o->x * vm[0] + o->fy * o->y * vm[1]) * rz2;
    const float* R = o->R;
    vp[0] += R[0] * v_x + R[3] * v_y + R[6] * v_z;
    vp[1] += R[1] * v_x + R[4] * v_y + R[7] * v_z;
    vp[2] += R[2] * v_x + R[5] * v_y + R[8] * v_z;
}

/* One iteration. grads: [N*11] = v_means(3) v_quats(4) v_scales(3) v_dens(1).
 * Returns the number of tile intersections. */
int64_t r2s_cpu_iteration(const float* means, const float* quats, const float* scales, const float* dens,
                          const float* views, const float* Ks, int N, int C, int H, int W, int T, float eps,
                          float clip, const float* target, float* image, float* grads, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    const int ntx = (W + T - 1) / T, nty = (H + T - 1) / T, nt = ntx * nty;
    float* covs = (float*)malloc(sizeof(float) * 6 * N);
    for (int n = 0; n < N; ++n) {
        float q[4], m, nn, R[9];
        quat_rot(quats + 4 * n, q, &m, &nn, R);
        const float* s = scales + 3 * n;
        int k = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = i; j < 3; ++j)
                covs[6 * n + k++] = R[3 * i] * R[3 * j] * s[0] * s[0] + R[3 * i + 1] * R[3 * j + 1] * s[1] * s[1] +
                                    R[3 * i + 2] * R[3 * j + 2] * s[2] * s[2];
    }
    float* part = (float*)calloc((size_t)C * N * 10, sizeof(float)); /* per-camera v_means(3) v_cov(6) v_dens */
    int64_t total_isects = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total_isects)
    for (int c = 0; c < C; ++c) {
        Proj* pr = (Proj*)malloc(sizeof(Proj) * N);
        int* cnt = (int*)calloc(nt + 1, sizeof(int));
        int* rect = (int*)malloc(sizeof(int) * 4 * N);
        for (int n = 0; n < N; ++n) {
            project_one(views + 16 * c, Ks + 9 * c, means + 3 * n, covs + 6 * n, W, H, eps, clip, &pr[n]);
            pr[n].opac = dens[n];
            int* r4 = rect + 4 * n;
            r4[0] = r4[1] = r4[2] = r4[3] = 0;
            if (!(pr[n].z > 0.01f)) continue;
            float a = pr[n].c2[0] + eps, d = pr[n].c2[3] + eps, b = pr[n].c2[1];
            if (!(a * d - b * b > 0)) continue;
            float rx = ceilf(clip * sqrtf(fmaxf(0.f, a))), ry = ceilf(clip * sqrtf(fmaxf(0.f, d)));
            if (!(rx > 0 && ry > 0) || !isfinite(rx) || !isfinite(ry)) continue;
            if (pr[n].mx + rx <= 0 || pr[n].mx - rx >= W || pr[n].my + ry <= 0 || pr[n].my - ry >= H) continue;
            if (span(pr[n].mx, (int)rx, T, ntx, &r4[0], &r4[1]) && span(pr[n].my, (int)ry, T, nty, &r4[2], &r4[3])) {
                for (int ty = r4[2]; ty < r4[3]; ++ty)
                    for (int tx = r4[0]; tx < r4[1]; ++tx) cnt[ty * ntx + tx + 1]++;
            } else {
                r4[0] = r4[1] = r4[2] = r4[3] = 0;
            }
        }
        for (int t = 0; t < nt; ++t) cnt[t + 1] += cnt[t];
        int n_is = cnt[nt];
        total_isects += n_is;
        int* list = (int*)malloc(sizeof(int) * (n_is > 0 ? n_is : 1));
        int* fill = (int*)malloc(sizeof(int) * nt);
        memcpy(fill, cnt, sizeof(int) * nt);
        for (int n = 0; n < N; ++n) { /* ascending n keeps each tile list stable */
            int* r4 = rect + 4 * n;
            for (int ty = r4[2]; ty < r4[3]; ++ty)
                for (int tx = r4[0]; tx < r4[1]; ++tx) list[fill[ty * ntx + tx]++] = n;
        }
        float* img = image + (size_t)c * H * W;
        float* vimg = (float*)malloc(sizeof(float) * H * W);
        /* forward */
        for (int t = 0; t < nt; ++t) {
            int ty = t / ntx, tx = t % ntx;
            for (int i = ty * T; i < ty * T + T && i < H; ++i)
                for (int j = tx * T; j < tx * T + T && j < W; ++j) {
                    float px = j + 0.5f, py = i + 0.5f, acc = 0.f;
                    for (int e = cnt[t]; e < cnt[t + 1]; ++e) {
                        const Proj* p = &pr[list[e]];
                        float dx = px - p->mx, dy = py - p->my;
                        float s = 0.5f * (p->a * dx * dx + p->c * dy * dy) + p->b * dx * dy;
                        float al = p->opac * p->mu * expf(-s);
                        if (s < 0 || fabsf(al) < 1e-5f) continue;
                        acc += al;
                    }
                    img[i * W + j] = acc;
                }
        }
        /* MSE adjoint */
        const float inv = 2.f / ((float)C * H * W);
        for (int k = 0; k < H * W; ++k) vimg[k] = inv * (img[k] - target[(size_t)c * H * W + k]);
        /* rasterize VJP, per (c, n) gradients of means2d, conic, opacity, mu */
        float* g7 = (float*)calloc((size_t)N * 7, sizeof(float));
        for (int t = 0; t < nt; ++t) {
            int ty = t / ntx, tx = t % ntx;
            for (int e = cnt[t]; e < cnt[t + 1]; ++e) {
                const int n = list[e];
                const Proj* p = &pr[n];
                float* g = g7 + 7 * n;
                for (int i = ty * T; i < ty * T + T && i < H; ++i)
                    for (int j = tx * T; j < tx * T + T && j < W; ++j) {
                        float dx = j + 0.5f - p->mx, dy = i + 0.5f - p->my;
                        float s = 0.5f * (p->a * dx * dx + p->c * dy * dy) + p->b * dx * dy;
                        float vis = expf(-s), al = p->opac * p->mu * vis;
                        if (s < 0 || fabsf(al) < 1e-5f) continue;
                        float v = vimg[i * W + j], vs = -v * al;
                        g[0] += -vs * (p->a * dx + p->b * dy);
                        g[1] += -vs * (p->c * dy + p->b * dx);
                        g[2] += vs * 0.5f * dx * dx;
                        g[3] += vs * dx * dy;
                        g[4] += vs * 0.5f * dy * dy;
                        g[5] += v * p->mu * vis;
                        g[6] += v * p->opac * vis;
                    }
            }
        }
        /* conic/mu and projection VJPs */
        for (int n = 0; n < N; ++n) {
            const Proj* p = &pr[n];
            float* out = part + ((size_t)c * N + n) * 10;
            const float* g = g7 + 7 * n;
            out[9] = g[5];
            if (!p->valid || !(p->z > 0.01f)) continue;
            double P0 = p->a, P1 = p->b, P2 = p->c, gv[6] = {0, 0, 0, 0, 0, 0};
            gv[0] = -(g[2] * P0 * P0 + g[3] * P0 * P1 + g[4] * P1 * P1);
            gv[1] = -(2 * g[2] * P0 * P1 + g[3] * (P0 * P2 + P1 * P1) + 2 * g[4] * P1 * P2);
            gv[3] = -(g[2] * P1 * P1 + g[3] * P1 * P2 + g[4] * P2 * P2);
            if (p->mu > 0) {
                double w0 = P0 * p->c2[2] + P1 * p->c2[4], w1 = P1 * p->c2[2] + P2 * p->c2[4];
                double vs = M_PI / p->mu * g[6];
                gv[0] += vs * w0 * w0; gv[1] += 2 * vs * w0 * w1; gv[2] -= 2 * vs * w0;
                gv[3] += vs * w1 * w1; gv[4] -= 2 * vs * w1; gv[5] += vs;
            }
            float vm[2] = {g[0], g[1]}, vc[6];
            for (int k = 0; k < 6; ++k) vc[k] = (float)gv[k];
            proj_vjp(p, vm, vc, out, out + 3);
        }
        free(g7); free(vimg); free(fill); free(list); free(rect); free(cnt); free(pr);
    }
    /* camera sum (fixed order) and parameter VJP */
    for (int n = 0; n < N; ++n) {
        float acc[10] = {0};
        for (int c = 0; c < C; ++c)
            for (int k = 0; k < 10; ++k) acc[k] += part[((size_t)c * N + n) * 10 + k];
        float q[4], m, nn, R[9];
        quat_rot(quats + 4 * n, q, &m, &nn, R);
        const float* s = scales + 3 * n;
        float* out = grads + 11 * (size_t)n;
        out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2];
        const float* vcv = acc + 3;
        float G[9] = {vcv[0], 0.5f * vcv[1], 0.5f * vcv[2], 0.5f * vcv[1], vcv[3], 0.5f * vcv[4],
                      0.5f * vcv[2], 0.5f * vcv[4], vcv[5]};
        float vR[9];
        for (int j = 0; j < 3; ++j) {
            float sg = 0;
            for (int i = 0; i < 3; ++i) {
                float gm = 2 * s[j] * (G[3 * i] * R[j] + G[3 * i + 1] * R[3 + j] + G[3 * i + 2] * R[6 + j]);
                vR[3 * i + j] = gm * s[j];
                sg += R[3 * i + j] * gm;
            }
            out[7 + j] = sg;
        }
        float w = q[0], x = q[1], y = q[2], z = q[3], gq[4];
        gq[0] = 2 * (x * (vR[7] - vR[5]) + y * (vR[2] - vR[6]) + z * (vR[3] - vR[1]));
        gq[1] = 2 * (-2 * x * (vR[4] + vR[8]) + y * (vR[1] + vR[3]) + z * (vR[2] + vR[6]) + w * (vR[7] - vR[5]));
        gq[2] = 2 * (x * (vR[1] + vR[3]) - 2 * y * (vR[0] + vR[8]) + z * (vR[5] + vR[7]) + w * (vR[2] - vR[6]));
        gq[3] = 2 * (x * (vR[2] + vR[6]) + y * (vR[5] + vR[7]) - 2 * z * (vR[0] + vR[4]) + w * (vR[3] - vR[1]));
        float qg = q[0] * gq[0] + q[1] * gq[1] + q[2] * gq[2] + q[3] * gq[3];
        for (int i = 0; i < 4; ++i) out[3 + i] = ((gq[i] - q[i] * qg) / nn) / m;
        out[10] = acc[9];
    }
    free(part);
    free(covs);
    return total_isects;
}

int r2s_cpu_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// axis.cpp -- per-axis operand matrices of the plane kernels (see AxisMats).
#include <cstring>
#include <vector>

#include "blockops.hpp"
#include "common.hpp"

namespace pisces {

namespace {

// round-to-nearest-even BF16 bit pattern of an FP32 value
uint16_t bf16_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf16_value(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void split(const std::vector<double> &a, std::vector<uint16_t> &hi, std::vector<uint16_t> &lo)
{
    hi.resize(a.size());
    lo.resize(a.size());
    for (size_t i = 0; i < a.size(); ++i) {
        const float x = (float)a[i];
        hi[i] = bf16_bits(x);
        lo[i] = bf16_bits(x - bf16_value(hi[i]));
    }
}

}  // namespace

void AxisStore::build(int n, const double *T, hipStream_t s)
{
    const int np = axis_pad(n);
    const size_t e2 = (size_t)np * np;
    std::vector<double> t(e2, 0.0), q(e2, 0.0), qt(e2, 0.0), lam(np, 0.0);
    std::vector<double> Q((size_t)n * n), W(n);
    // eigenvectors of T (columns of Q)
    syev_host(n, T, n, W.data(), Q.data(), n);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) {
            t[(size_t)r * np + c] = T[(size_t)r * n + c];
            // syev_host returns column-major eigenvectors: Q(r, c) = Q[r + n c]
            q[(size_t)r * np + c] = Q[(size_t)r + (size_t)n * c];
            qt[(size_t)c * np + r] = Q[(size_t)r + (size_t)n * c];
        }
    for (int i = 0; i < n; ++i) lam[i] = W[i];
    std::vector<float> t32(e2);
    for (size_t i = 0; i < e2; ++i) t32[i] = (float)t[i];
    std::vector<uint16_t> th, tl, qh, ql, qth, qtl;
    split(t, th, tl);
    split(q, qh, ql);
    split(qt, qth, qtl);
    // one allocation: FP64 arrays first (t, q, qt, lam), then FP32, then BF16,
    // each array starting on a 256-byte boundary
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t s64 = up(8 * e2), slam = up(8 * (size_t)np), s32 = up(4 * e2), s16 = up(2 * e2);
    const size_t total = 3 * s64 + slam + s32 + 6 * s16;
    std::vector<unsigned char> host(total, 0);
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes, size_t stride) {
        std::memcpy(host.data() + off, src, bytes);
        const size_t o = off;
        off += stride;
        return o;
    };
    const size_t o_t = put(t.data(), 8 * e2, s64), o_q = put(q.data(), 8 * e2, s64),
                 o_qt = put(qt.data(), 8 * e2, s64), o_lam = put(lam.data(), 8 * (size_t)np, slam),
                 o_t32 = put(t32.data(), 4 * e2, s32), o_th = put(th.data(), 2 * e2, s16),
                 o_tl = put(tl.data(), 2 * e2, s16), o_qh = put(qh.data(), 2 * e2, s16),
                 o_ql = put(ql.data(), 2 * e2, s16), o_qth = put(qth.data(), 2 * e2, s16),
                 o_qtl = put(qtl.data(), 2 * e2, s16);
    buf.alloc(total);
    PC_HIP(hipMemcpyAsync(buf.p, host.data(), total, hipMemcpyHostToDevice, s));
    PC_HIP(hipStreamSynchronize(s));
    const unsigned char *b = buf.as<unsigned char>();
    m.n = n;
    m.np = np;
    m.t64 = reinterpret_cast<const double *>(b + o_t);
    m.q64 = reinterpret_cast<const double *>(b + o_q);
    m.qt64 = reinterpret_cast<const double *>(b + o_qt);
    m.lam = reinterpret_cast<const double *>(b + o_lam);
    m.t32 = reinterpret_cast<const float *>(b + o_t32);
    m.th = reinterpret_cast<const uint16_t *>(b + o_th);
    m.tl = reinterpret_cast<const uint16_t *>(b + o_tl);
    m.qh = reinterpret_cast<const uint16_t *>(b + o_qh);
    m.ql = reinterpret_cast<const uint16_t *>(b + o_ql);
    m.qth = reinterpret_cast<const uint16_t *>(b + o_qth);
    m.qtl = reinterpret_cast<const uint16_t *>(b + o_qtl);
}

}  // namespace pisces

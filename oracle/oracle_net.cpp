// CPU oracle, part 4: the learned error-correction network (no reference
// counterpart; the reference has no learned component).
//
// TEST INFRASTRUCTURE ONLY: the scalar statement the GPU network kernels are
// checked against.  Pinned in tests by central finite differences of the FP64
// mode, by torch.nn.functional.conv3d on the CPU, and by closed forms (a
// single-tap identity network reproduces its input).
//
// Network (per 16^3 block): input X0 = 20^3 x 2 channels (the block with a
// 2-cell halo: scaled residual and phase field), four 3^3 convolutions with 16
// hidden channels, ReLU after the first three:
//   L1 valid 20^3 x 2  -> 18^3 x 16      L2 valid 18^3 -> 16^3 x 16
//   L3 same  16^3 x 16 -> 16^3 x 16      L4 same  16^3 x 16 -> 16^3 x 1
// ("same" = zero padding of one cell).  out[p][o] = bias[o] +
// sum_t sum_c W[t][c][o] in[p + off + d_t][c], tap t = dx + 3 dy + 9 dz with
// d_t = (dx, dy, dz) in {0,1,2}^3, off = 0 (valid) or -1 (same).
// Parameters are stored padded: per layer W[27][16][16] (c < cin, o < cout
// used) + bias[16]; layer l at offset l * 6928.
// Tensors are channels-last: [z][y][x][c].
//
// mode 1 ("bf16") rounds where the GPU stores bf16: X0, the weights, the
// activations A1..A3, and the backward's gradient tensors G4..G1; sums in
// FP64.  mode 0 is plain FP64.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

typedef int64_t i64;

namespace {

constexpr int kLayerParams = 27 * 16 * 16 + 16;

double bf16r(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return v;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

struct Layer {
    int sin, sout, off, cin, cout;
    bool relu;
};
const Layer kLayers[4] = {{20, 18, 0, 2, 16, true}, {18, 16, 0, 16, 16, true}, {16, 16, -1, 16, 16, true},
                          {16, 16, -1, 16, 1, false}};

struct Net {
    std::vector<double> w;  // padded params, rounded if mode 1
    int mode;
    double rnd(double v) const { return mode ? bf16r(v) : v; }
    const double *W(int l) const { return &w[(size_t)l * kLayerParams]; }
    const double *B(int l) const { return &w[(size_t)l * kLayerParams + 27 * 256]; }
};

// out[p][o] (sout^3 x 16, o < cout) = bias + conv(in)
void conv_fwd(const Net &net, int l, const std::vector<double> &in, std::vector<double> &out) {
    const Layer &L = kLayers[l];
    const double *W = net.W(l), *Bs = net.B(l);
    out.assign((size_t)L.sout * L.sout * L.sout * 16, 0.0);
    for (int z = 0; z < L.sout; ++z)
        for (int y = 0; y < L.sout; ++y)
            for (int x = 0; x < L.sout; ++x) {
                double acc[16];
                for (int o = 0; o < 16; ++o) acc[o] = o < L.cout ? Bs[o] : 0.0;
                for (int t = 0; t < 27; ++t) {
                    const int ix = x + L.off + t % 3, iy = y + L.off + (t / 3) % 3, iz = z + L.off + t / 9;
                    if (ix < 0 || iy < 0 || iz < 0 || ix >= L.sin || iy >= L.sin || iz >= L.sin) continue;
                    const double *a = &in[(((size_t)iz * L.sin + iy) * L.sin + ix) * 16];
                    for (int c = 0; c < L.cin; ++c)
                        for (int o = 0; o < L.cout; ++o) acc[o] += W[(t * 16 + c) * 16 + o] * a[c];
                }
                double *q = &out[(((size_t)z * L.sout + y) * L.sout + x) * 16];
                for (int o = 0; o < L.cout; ++o) q[o] = L.relu ? (acc[o] > 0.0 ? acc[o] : 0.0) : acc[o];
            }
}

// Forward of one tile; acts[0] = X0 (20^3 x 16, channels >= 2 zero), acts[1..3]
// = A1..A3, acts[4] = Y (16^3 x 16, channel 0).
void forward(const Net &net, const float *x0, std::vector<double> acts[5]) {
    acts[0].assign((size_t)8000 * 16, 0.0);
    for (int i = 0; i < 8000; ++i)
        for (int c = 0; c < 2; ++c) acts[0][(size_t)i * 16 + c] = net.rnd(x0[(size_t)i * 2 + c]);
    for (int l = 0; l < 4; ++l) {
        conv_fwd(net, l, acts[l], acts[l + 1]);
        if (l < 3)
            for (double &v : acts[l + 1]) v = net.rnd(v);
    }
}

Net make_net(const float *params, int mode) {
    Net n;
    n.mode = mode;
    n.w.assign((size_t)4 * kLayerParams, 0.0);
    for (int l = 0; l < 4; ++l) {
        const Layer &L = kLayers[l];
        for (int t = 0; t < 27; ++t)
            for (int c = 0; c < L.cin; ++c)
                for (int o = 0; o < L.cout; ++o) {
                    const size_t k = (size_t)l * kLayerParams + (t * 16 + c) * 16 + o;
                    n.w[k] = n.rnd(params[k]);
                }
        for (int o = 0; o < L.cout; ++o) {
            const size_t k = (size_t)l * kLayerParams + 27 * 256 + o;
            n.w[k] = params[k];  // biases stay FP32 on the GPU
        }
    }
    return n;
}

}  // namespace

extern "C" {

int orc_net_param_count() { return 4 * kLayerParams; }

// y[t][16^3] = network output for n tiles x0[t][20^3][2].
void orc_net_forward(const float *params, const float *x0, i64 n, int mode, float *y) {
    const Net net = make_net(params, mode);
#pragma omp parallel for schedule(dynamic)
    for (i64 t = 0; t < n; ++t) {
        std::vector<double> acts[5];
        forward(net, x0 + (size_t)t * 16000, acts);
        for (int i = 0; i < 4096; ++i) y[(size_t)t * 4096 + i] = (float)acts[4][(size_t)i * 16];
    }
}

// Loss L = sum over tiles and cells of (y - target)^2 / (2 n 4096) and its
// gradient with respect to the padded parameters (grad has orc_net_param_count
// entries; padding entries are zero).  Returns L.
double orc_net_loss_grad(const float *params, const float *x0, const float *target, i64 n, int mode,
                         double *grad) {
    const Net net = make_net(params, mode);
    const int P = 4 * kLayerParams;
    std::memset(grad, 0, sizeof(double) * P);
    const double inv = 1.0 / ((double)n * 4096.0);
    double loss = 0.0;
    for (i64 tile = 0; tile < n; ++tile) {
        std::vector<double> acts[5];
        forward(net, x0 + (size_t)tile * 16000, acts);
        // G = dL / d(pre-activation output) of each layer, channels-last x 16
        std::vector<double> G((size_t)4096 * 16, 0.0);
        for (int i = 0; i < 4096; ++i) {
            const double d = acts[4][(size_t)i * 16] - target[(size_t)tile * 4096 + i];
            loss += 0.5 * d * d * inv;
            G[(size_t)i * 16] = net.rnd(d * inv);
        }
        for (int l = 3; l >= 0; --l) {
            const Layer &L = kLayers[l];
            const std::vector<double> &X = acts[l];
            double *gW = grad + (size_t)l * kLayerParams, *gB = gW + 27 * 256;
            // weight and bias gradients
            for (int z = 0; z < L.sout; ++z)
                for (int y = 0; y < L.sout; ++y)
                    for (int x = 0; x < L.sout; ++x) {
                        const double *g = &G[(((size_t)z * L.sout + y) * L.sout + x) * 16];
                        for (int o = 0; o < L.cout; ++o) gB[o] += g[o];
                        for (int t = 0; t < 27; ++t) {
                            const int ix = x + L.off + t % 3, iy = y + L.off + (t / 3) % 3, iz = z + L.off + t / 9;
                            if (ix < 0 || iy < 0 || iz < 0 || ix >= L.sin || iy >= L.sin || iz >= L.sin) continue;
                            const double *a = &X[(((size_t)iz * L.sin + iy) * L.sin + ix) * 16];
                            for (int c = 0; c < L.cin; ++c)
                                for (int o = 0; o < L.cout; ++o) gW[(t * 16 + c) * 16 + o] += a[c] * g[o];
                        }
                    }
            if (l == 0) break;
            // gradient with respect to the input activation, masked by the
            // previous ReLU (A > 0)
            std::vector<double> Gin((size_t)L.sin * L.sin * L.sin * 16, 0.0);
            const double *W = net.W(l);
            for (int z = 0; z < L.sout; ++z)
                for (int y = 0; y < L.sout; ++y)
                    for (int x = 0; x < L.sout; ++x) {
                        const double *g = &G[(((size_t)z * L.sout + y) * L.sout + x) * 16];
                        for (int t = 0; t < 27; ++t) {
                            const int ix = x + L.off + t % 3, iy = y + L.off + (t / 3) % 3, iz = z + L.off + t / 9;
                            if (ix < 0 || iy < 0 || iz < 0 || ix >= L.sin || iy >= L.sin || iz >= L.sin) continue;
                            double *gi = &Gin[(((size_t)iz * L.sin + iy) * L.sin + ix) * 16];
                            for (int c = 0; c < L.cin; ++c)
                                for (int o = 0; o < L.cout; ++o) gi[c] += W[(t * 16 + c) * 16 + o] * g[o];
                        }
                    }
            for (size_t i = 0; i < Gin.size(); ++i) Gin[i] = X[i] > 0.0 ? net.rnd(Gin[i]) : 0.0;
            G.swap(Gin);
        }
    }
    return loss;
}

}  // extern "C"

extern "C" {

// Network input of the listed blocks of a band (x0: n x 20^3 x 2, bf16 bits):
// halo cell h of block k is the cell origin + h - 2, mirrored into the domain;
// channel 0 = bf16(r * inv) if its block is listed else 0, channel 1 =
// bf16(phi) read through the phase field grid's own bid -> slot table.
void orc_net_input(i64 sx, i64 sy, i64 sz, const int32_t *b2s, const i64 *active, i64 n, const float *r, float inv,
                   const float *phi_slab, const int32_t *phi_b2s, uint16_t *x0) {
    const i64 nx = (sx + 15) >> 4, ny = (sy + 15) >> 4, nb = nx * ny * ((sz + 15) >> 4);
    std::vector<char> listed((size_t)nb, 0);
    for (i64 k = 0; k < n; ++k) listed[active[k]] = 1;
    auto mir = [](i64 g, i64 s) { return g < 0 ? -g - 1 : (g >= s ? 2 * s - 1 - g : g); };
    auto bits = [](float v) {
        const double b = bf16r(v);
        const float f = (float)b;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return (uint16_t)(u >> 16);
    };
    for (i64 k = 0; k < n; ++k) {
        const i64 bid0 = active[k];
        const i64 ox = (bid0 % nx) * 16, oy = ((bid0 / nx) % ny) * 16, oz = (bid0 / (nx * ny)) * 16;
        for (int c = 0; c < 8000; ++c) {
            const i64 gx = mir(ox + c % 20 - 2, sx), gy = mir(oy + (c / 20) % 20 - 2, sy), gz = mir(oz + c / 400 - 2, sz);
            const i64 bid = (gx >> 4) + (gy >> 4) * nx + (gz >> 4) * nx * ny;
            const i64 vid = (gx & 15) | ((gy & 15) << 4) | ((gz & 15) << 8);
            const float rv = listed[bid] ? r[(i64)b2s[bid] * 4096 + vid] * inv : 0.0f;
            x0[((size_t)k * 8000 + c) * 2] = bits(rv);
            x0[((size_t)k * 8000 + c) * 2 + 1] = bits(phi_slab[(i64)phi_b2s[bid] * 4096 + vid]);
        }
    }
}

}  // extern "C"

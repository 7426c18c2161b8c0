// learn.hip -- learned starting values for the model fit: a GPU forward
// simulator fused with site-pattern / pairwise-distance feature extraction,
// and a small MLP (FP32) trained with hand-written forward / backward
// kernels.  The simulator extends simpleML::simulate_chars
// (simpleML.cpp:233-292) with gamma rate categories, gaps and a seeded
// counter-based RNG.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/phylo_hip.h"
#include "../../include/phylo_learn.h"

namespace {

constexpr int kSimThreads = 256;
constexpr int kMaxSimTaxa = 32;
constexpr int kMaxSimNodes = 2 * kMaxSimTaxa - 1;
constexpr int kMaxSimCat = 8;
constexpr int kFeat = 256;
constexpr unsigned char kGap = 255;

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// uniform in [0, 1) from a counter
__device__ inline float urand(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    const uint64_t h = mix64(seed ^ mix64(a * 0x9e3779b97f4a7c15ull ^ mix64(b * 0xc2b2ae3d27d4eb4full ^ (c + 1))));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// One workgroup per alignment.  Per-alignment inputs (eig = V[16], Vinv[16],
// lam[4]; pi[4]; cat_rates[C]) give P(t) = V exp(lam r t) Vinv in LDS as
// cumulative rows; each thread simulates one site of a 256-site chunk down
// the tree in preorder, then the chunk's tip states update the features.
__global__ __launch_bounds__(kSimThreads) void sim_features_kernel(
    int n_taxa, int n_sites, int n_cat, const int* __restrict__ left, const int* __restrict__ right,
    const int* __restrict__ roots, const double* __restrict__ blen, const double* __restrict__ eig,
    const double* __restrict__ pi_all, const double* __restrict__ cat_all, float gap_frac, uint64_t seed,
    float* __restrict__ feat, uint32_t* __restrict__ masks_out) {
    const int a = blockIdx.x;
    const int tid = threadIdx.x;
    const int n_nodes = 2 * n_taxa - 1;
    __shared__ float cumP[kMaxSimNodes][kMaxSimCat][16];
    __shared__ unsigned char st[kMaxSimNodes][kSimThreads];
    __shared__ int pre[kMaxSimNodes];
    __shared__ int par[kMaxSimNodes];
    __shared__ float cum_pi[4];
    __shared__ int h_site[4 + 32 + 32];
    __shared__ float h_pair[16 + 32 + 32 + 32 + 8];
    __shared__ int base[4];
    __shared__ int gaps_total;
    __shared__ int invariant;
    const int* L = left + (size_t)a * n_nodes;
    const int* R = right + (size_t)a * n_nodes;
    const double* bl = blen + (size_t)a * n_nodes;
    const double* V = eig + (size_t)a * 36;
    const double* Vi = V + 16;
    const double* lam = V + 32;
    for (int i = tid; i < 4 + 32 + 32; i += kSimThreads) h_site[i] = 0;
    for (int i = tid; i < 16 + 32 + 32 + 32 + 8; i += kSimThreads) h_pair[i] = 0.f;
    if (tid < 4) base[tid] = 0;
    if (tid == 0) {
        gaps_total = 0;
        invariant = 0;
        // preorder by an explicit stack, parents recorded
        int stack[kMaxSimNodes];
        int sp = 0, k = 0;
        stack[sp++] = roots[a];
        par[roots[a]] = -1;
        while (sp) {
            const int v = stack[--sp];
            pre[k++] = v;
            if (R[v] >= 0) { par[R[v]] = v; stack[sp++] = R[v]; }
            if (L[v] >= 0) { par[L[v]] = v; stack[sp++] = L[v]; }
        }
        double c = 0.0;
        for (int i = 0; i < 4; ++i) { c += pi_all[(size_t)a * 4 + i]; cum_pi[i] = (float)c; }
        cum_pi[3] = 1.0f;
    }
    // cumulative P rows: cumP[v][c][i*4 + j] = sum_{j' <= j} P_ij'
    for (int e = tid; e < n_nodes * n_cat * 4; e += kSimThreads) {
        const int i = e % 4, c = (e / 4) % n_cat, v = e / (4 * n_cat);
        const double t = bl[v] * cat_all[(size_t)a * n_cat + c];
        double acc = 0.0;
        for (int j = 0; j < 4; ++j) {
            double p = 0.0;
            for (int k = 0; k < 4; ++k) p += V[i * 4 + k] * exp(lam[k] * t) * Vi[k * 4 + j];
            acc += p > 0.0 ? p : 0.0;
            cumP[v][c][i * 4 + j] = (float)acc;
        }
        const float tot = cumP[v][c][i * 4 + 3];
        for (int j = 0; j < 4; ++j) cumP[v][c][i * 4 + j] /= tot;
        cumP[v][c][i * 4 + 3] = 1.0f;
    }
    __syncthreads();
    // pair ownership: thread owns pairs p = tid, tid + 256 (496 pairs for 32 taxa)
    const int n_pairs = n_taxa * (n_taxa - 1) / 2;
    int pa[2] = {-1, -1}, pb[2] = {-1, -1};
    for (int q = 0; q < 2; ++q) {
        int p = tid + q * kSimThreads;
        if (p >= n_pairs) continue;
        int i = 0;
        while (p >= n_taxa - 1 - i) { p -= n_taxa - 1 - i; ++i; }
        pa[q] = i;
        pb[q] = i + 1 + p;
    }
    int joint[2][16];
    for (int q = 0; q < 2; ++q)
        for (int c = 0; c < 16; ++c) joint[q][c] = 0;
    int basec[4] = {0, 0, 0, 0}, gapc = 0;
    for (int s0 = 0; s0 < n_sites; s0 += kSimThreads) {
        const int s = s0 + tid;
        if (s < n_sites) {
            const int cat = n_cat > 1 ? min(n_cat - 1, (int)(urand(seed, a, s, 1u << 20) * n_cat)) : 0;
            const float u = urand(seed, a, s, (1u << 20) + 1);
            int rs = 3;
            for (int j = 0; j < 4; ++j)
                if (u < cum_pi[j]) { rs = j; break; }
            st[pre[0]][tid] = (unsigned char)rs;
            for (int k = 1; k < n_nodes; ++k) {
                const int v = pre[k];
                const int anc = st[par[v]][tid];
                const float r = urand(seed, a, s, (uint64_t)v);
                const float* row = &cumP[v][cat][anc * 4];
                int x = 3;
                for (int j = 0; j < 4; ++j)
                    if (r < row[j]) { x = j; break; }
                st[v][tid] = (unsigned char)x;
            }
            // gaps at the tips, site summaries
            int cnt[4] = {0, 0, 0, 0}, n_ok = 0;
            for (int t = 0; t < n_taxa; ++t) {
                if (gap_frac > 0.f && urand(seed, a, s, (2u << 20) + t) < gap_frac) {
                    st[t][tid] = kGap;
                    ++gapc;
                } else {
                    ++cnt[st[t][tid]];
                    ++basec[st[t][tid]];
                    ++n_ok;
                }
                if (masks_out)
                    masks_out[((size_t)a * n_taxa + t) * n_sites + s] = st[t][tid] == kGap ? 15u : (1u << st[t][tid]);
            }
            int distinct = 0, mx = 0, sq = 0;
            for (int j = 0; j < 4; ++j) {
                distinct += cnt[j] > 0;
                mx = max(mx, cnt[j]);
                sq += cnt[j] * cnt[j];
            }
            if (distinct > 0) atomicAdd(&h_site[distinct - 1], 1);
            if (distinct == 1) atomicAdd(&invariant, 1);
            // minor-allele fraction (n - max) / n in bins of 1/64 and
            // Gini-Simpson diversity 1 - sum f^2 in bins of 3/128, both from
            // integers so the bin edges are exact
            const int nn2 = n_ok * n_ok;
            atomicAdd(&h_site[4 + (n_ok ? min(31, 64 * (n_ok - mx) / n_ok) : 0)], 1);
            atomicAdd(&h_site[36 + (n_ok ? min(31, 128 * (nn2 - sq) / (3 * nn2)) : 0)], 1);
        }
        __syncthreads();
        const int len = min(kSimThreads, n_sites - s0);
        for (int q = 0; q < 2; ++q) {
            if (pa[q] < 0) continue;
            const unsigned char* x = st[pa[q]];
            const unsigned char* y = st[pb[q]];
            for (int k = 0; k < len; ++k)
                if (x[k] != kGap && y[k] != kGap) ++joint[q][x[k] * 4 + y[k]];
        }
        __syncthreads();
    }
    // pair statistics
    for (int q = 0; q < 2; ++q) {
        if (pa[q] < 0) continue;
        int n = 0;
        for (int c = 0; c < 16; ++c) n += joint[q][c];
        if (n == 0) continue;
        const float inv = 1.f / n;
        const int same = joint[q][0] + joint[q][5] + joint[q][10] + joint[q][15];
        const int ts = joint[q][1] + joint[q][4] + joint[q][11] + joint[q][14];     // A<->G, C<->T
        const int tv = n - same - ts;
        const float pd = (n - same) * inv, pts = ts * inv, ptv = tv * inv;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) atomicAdd(&h_pair[i * 4 + j], 0.5f * (joint[q][i * 4 + j] + joint[q][j * 4 + i]) * inv / n_pairs);
        // p-distance in bins of 1/40, transition and transversion distances
        // in bins of 1/80 (integer bin edges)
        atomicAdd(&h_pair[16 + min(31, 40 * (n - same) / n)], 1.f / n_pairs);
        atomicAdd(&h_pair[48 + min(31, 80 * ts / n)], 1.f / n_pairs);
        atomicAdd(&h_pair[80 + min(31, 80 * tv / n)], 1.f / n_pairs);
        atomicAdd(&h_pair[112], pd / n_pairs);
        atomicAdd(&h_pair[113], pd * pd / n_pairs);
        atomicAdd(&h_pair[114], pts / n_pairs);
        atomicAdd(&h_pair[115], ptv / n_pairs);
        // Jukes-Cantor distance capped at 3
        const float jc = fminf(3.f, -0.75f * logf(fmaxf(1e-6f, 1.f - pd / 0.75f)));
        atomicAdd(&h_pair[116], jc / n_pairs);
    }
    for (int j = 0; j < 4; ++j) atomicAdd(&base[j], basec[j]);
    atomicAdd(&gaps_total, gapc);
    __syncthreads();
    float* f = feat + (size_t)a * kFeat;
    for (int i = tid; i < kFeat; i += kSimThreads) {
        float v = 0.f;
        const int nb = base[0] + base[1] + base[2] + base[3];
        if (i < 4) v = nb ? (float)base[i] / nb : 0.f;
        else if (i < 20) v = h_pair[i - 4];
        else if (i < 116) v = h_pair[16 + (i - 20)];
        else if (i < 120) v = (float)h_site[i - 116] / n_sites;
        else if (i < 184) v = (float)h_site[4 + (i - 120)] / n_sites;
        else if (i == 184) v = h_pair[112];
        else if (i == 185) v = sqrtf(fmaxf(0.f, h_pair[113] - h_pair[112] * h_pair[112]));
        else if (i == 186) v = h_pair[114];
        else if (i == 187) v = h_pair[115];
        else if (i == 188) v = (float)gaps_total / ((float)n_sites * n_taxa);
        else if (i == 189) v = h_pair[116];
        else if (i == 190) v = fminf(20.f, h_pair[114] / fmaxf(1e-6f, h_pair[115]));
        else if (i == 191) v = (float)invariant / n_sites;
        f[i] = v;
    }
}

// ---------------------------------------------------------------------------
// MLP kernels (FP32)
// ---------------------------------------------------------------------------

constexpr int kTile = 64;

// C[M x N] (+)= op(A) op(B); op(A) is M x K (A stored K x M when ta), op(B)
// is K x N (B stored N x K when tb).  64 x 64 tiles through LDS, 256 threads,
// 4 x 4 outputs per thread.
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                   const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, int ta, int tb, int accumulate) {
    __shared__ float As[16][kTile + 4];
    __shared__ float Bs[16][kTile + 4];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * kTile, col0 = blockIdx.x * kTile;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * kTile; e += 256) {
            const int kk = e / kTile, mm = e % kTile;
            const int r = row0 + mm, k = k0 + kk;
            float va = 0.f;
            if (r < M && k < K) va = ta ? A[(size_t)k * lda + r] : A[(size_t)r * lda + k];
            As[kk][mm] = va;
            const int c = col0 + mm;
            float vb = 0.f;
            if (c < N && k < K) vb = tb ? B[(size_t)c * ldb + k] : B[(size_t)k * ldb + c];
            Bs[kk][mm] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int r = row0 + ty * 4 + i, c = col0 + tx * 4 + j;
            if (r < M && c < N) {
                float* dst = C + (size_t)r * ldc + c;
                *dst = accumulate ? *dst + acc[i][j] : acc[i][j];
            }
        }
}

// Y = X + bias (per column), optionally ReLU.
__global__ void bias_act_kernel(int M, int N, float* __restrict__ Y, const float* __restrict__ b, int relu) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * N) return;
    float v = Y[i] + b[i % N];
    Y[i] = relu ? fmaxf(v, 0.f) : v;
}

// dY = 2 (Y - T) / (M N); loss partial sums per block.
__global__ void mse_grad_kernel(int M, int N, const float* __restrict__ Y, const float* __restrict__ T,
                                float* __restrict__ dY, float* __restrict__ loss_part) {
    __shared__ float red[256];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < (long)M * N) {
        const float d = Y[i] - T[i];
        dY[i] = 2.f * d / ((float)M * N);
        l = d * d / ((float)M * N);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ *= (H > 0): the ReLU gradient through the stored activation.
__global__ void relu_grad_kernel(long n, const float* __restrict__ H, float* __restrict__ dZ) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(H[i] > 0.f)) dZ[i] = 0.f;
}

// db[j] = sum_i dZ[i][j]
__global__ void col_sum_kernel(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int i = 0; i < M; ++i) s += dZ[(size_t)i * N + j];
    db[j] = s;
}

__global__ void adam_kernel(long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void gather_rows_kernel(int rows, int cols, const int* __restrict__ idx, const float* __restrict__ src,
                                   float* __restrict__ dst) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * cols) return;
    const int r = i / cols, c = i % cols;
    dst[i] = src[(size_t)idx[r] * cols + c];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct phylo_mlp {
    int device = 0;
    int n_in = 0, n_hid = 0, n_out = 0, max_batch = 0;
    long n_params = 0;
    hipStream_t stream = nullptr;
    std::string err;
    float* d_w = nullptr;       // [W1 | b1 | W2 | b2 | W3 | b3]
    float* d_g = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_x = nullptr;       // batch buffers
    float* d_t = nullptr;
    float* d_h1 = nullptr;
    float* d_h2 = nullptr;
    float* d_y = nullptr;
    float* d_dh1 = nullptr;
    float* d_dh2 = nullptr;
    float* d_dy = nullptr;
    float* d_loss = nullptr;
    int* d_idx = nullptr;
    float* d_data_x = nullptr;  // training set
    float* d_data_t = nullptr;
    int n_data = 0;
    long step = 0;
};

namespace {

int mfail(phylo_mlp* p, int code, const std::string& m) {
    if (p) p->err = m;
    return code;
}
#define MCHK(p, expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return mfail((p), PHYLO_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

void gemm(phylo_mlp* p, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
          int ta, int tb, int acc) {
    dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
    hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, p->stream, M, N, K, A, lda, B, ldb, C, ldc, ta, tb, acc);
}

inline unsigned nblk(long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Parameter views inside the flat buffer.
struct Views {
    float *W1, *b1, *W2, *b2, *W3, *b3;
};
Views views(phylo_mlp* p, float* base) {
    Views v;
    v.W1 = base;
    v.b1 = v.W1 + (size_t)p->n_in * p->n_hid;
    v.W2 = v.b1 + p->n_hid;
    v.b2 = v.W2 + (size_t)p->n_hid * p->n_hid;
    v.W3 = v.b2 + p->n_hid;
    v.b3 = v.W3 + (size_t)p->n_hid * p->n_out;
    return v;
}

void forward(phylo_mlp* p, int B, const float* X) {
    Views w = views(p, p->d_w);
    gemm(p, B, p->n_hid, p->n_in, X, p->n_in, w.W1, p->n_hid, p->d_h1, p->n_hid, 0, 0, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((long)B * p->n_hid)), dim3(256), 0, p->stream, B, p->n_hid, p->d_h1, w.b1, 1);
    gemm(p, B, p->n_hid, p->n_hid, p->d_h1, p->n_hid, w.W2, p->n_hid, p->d_h2, p->n_hid, 0, 0, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((long)B * p->n_hid)), dim3(256), 0, p->stream, B, p->n_hid, p->d_h2, w.b2, 1);
    gemm(p, B, p->n_out, p->n_hid, p->d_h2, p->n_hid, w.W3, p->n_out, p->d_y, p->n_out, 0, 0, 0);
    hipLaunchKernelGGL(bias_act_kernel, dim3(nblk((long)B * p->n_out)), dim3(256), 0, p->stream, B, p->n_out, p->d_y, w.b3, 0);
}

// Gradient of the mean squared error into d_g; loss partials in d_loss.
void backward(phylo_mlp* p, int B, const float* X, const float* T) {
    Views w = views(p, p->d_w), g = views(p, p->d_g);
    const long ny = (long)B * p->n_out;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(nblk(ny)), dim3(256), 0, p->stream, B, p->n_out, p->d_y, T, p->d_dy, p->d_loss);
    // layer 3
    gemm(p, p->n_hid, p->n_out, B, p->d_h2, p->n_hid, p->d_dy, p->n_out, g.W3, p->n_out, 1, 0, 0);
    hipLaunchKernelGGL(col_sum_kernel, dim3(nblk(p->n_out)), dim3(256), 0, p->stream, B, p->n_out, p->d_dy, g.b3);
    gemm(p, B, p->n_hid, p->n_out, p->d_dy, p->n_out, w.W3, p->n_out, p->d_dh2, p->n_hid, 0, 1, 0);
    hipLaunchKernelGGL(relu_grad_kernel, dim3(nblk((long)B * p->n_hid)), dim3(256), 0, p->stream, (long)B * p->n_hid, p->d_h2, p->d_dh2);
    // layer 2
    gemm(p, p->n_hid, p->n_hid, B, p->d_h1, p->n_hid, p->d_dh2, p->n_hid, g.W2, p->n_hid, 1, 0, 0);
    hipLaunchKernelGGL(col_sum_kernel, dim3(nblk(p->n_hid)), dim3(256), 0, p->stream, B, p->n_hid, p->d_dh2, g.b2);
    gemm(p, B, p->n_hid, p->n_hid, p->d_dh2, p->n_hid, w.W2, p->n_hid, p->d_dh1, p->n_hid, 0, 1, 0);
    hipLaunchKernelGGL(relu_grad_kernel, dim3(nblk((long)B * p->n_hid)), dim3(256), 0, p->stream, (long)B * p->n_hid, p->d_h1, p->d_dh1);
    // layer 1
    gemm(p, p->n_in, p->n_hid, B, X, p->n_in, p->d_dh1, p->n_hid, g.W1, p->n_hid, 1, 0, 0);
    hipLaunchKernelGGL(col_sum_kernel, dim3(nblk(p->n_hid)), dim3(256), 0, p->stream, B, p->n_hid, p->d_dh1, g.b1);
}

float read_loss(phylo_mlp* p, int B) {
    const long ny = (long)B * p->n_out;
    std::vector<float> part(nblk(ny));
    if (hipMemcpyAsync(part.data(), p->d_loss, part.size() * sizeof(float), hipMemcpyDeviceToHost, p->stream) !=
            hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
        return NAN;
    double s = 0.0;
    for (float x : part) s += x;
    return (float)s;
}

}  // namespace

extern "C" {

int phylo_sim_features(int device, int n_align, int n_taxa, int n_sites, int n_cat, const int32_t* left,
                       const int32_t* right, const int32_t* roots, const double* blen, const double* eig,
                       const double* pi, const double* cat_rates, double gap_frac, uint64_t seed, float* features,
                       uint32_t* masks) {
    if (n_align < 1 || n_taxa < 4 || n_taxa > kMaxSimTaxa || n_sites < 1 || n_sites > (1 << 20) || n_cat < 1 ||
        n_cat > kMaxSimCat ||
        !left || !right || !roots || !blen || !eig || !pi || !cat_rates || !features || gap_frac < 0 || gap_frac >= 1)
        return PHYLO_ERR_ARG;
    const int nn = 2 * n_taxa - 1;
    for (long i = 0; i < (long)n_align * nn; ++i)
        if (left[i] < -1 || left[i] >= nn || right[i] < -1 || right[i] >= nn) return PHYLO_ERR_ARG;
    // every tree must be a rooted binary tree over nodes 0..nn-1 with the tips
    // 0..n_taxa-1 childless: the kernel's preorder stack and state table are
    // sized by that
    std::vector<int> n_par(nn);
    for (int k = 0; k < n_align; ++k) {
        const int32_t* L = left + (size_t)k * nn;
        const int32_t* R = right + (size_t)k * nn;
        if (roots[k] < n_taxa || roots[k] >= nn) return PHYLO_ERR_ARG;
        std::fill(n_par.begin(), n_par.end(), 0);
        for (int v = 0; v < nn; ++v) {
            const bool tip = v < n_taxa;
            if (tip != (L[v] < 0) || tip != (R[v] < 0)) return PHYLO_ERR_ARG;
            if (!tip) {
                ++n_par[L[v]];
                ++n_par[R[v]];
            }
        }
        for (int v = 0; v < nn; ++v)
            if (n_par[v] != (v == roots[k] ? 0 : 1)) return PHYLO_ERR_ARG;
        // connected: count nodes reachable from the root
        std::vector<int> stack{roots[k]};
        int seen = 0;
        while (!stack.empty()) {
            const int v = stack.back();
            stack.pop_back();
            if (++seen > nn) return PHYLO_ERR_ARG;
            if (L[v] >= 0) { stack.push_back(L[v]); stack.push_back(R[v]); }
        }
        if (seen != nn) return PHYLO_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) return PHYLO_ERR_HIP;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return PHYLO_ERR_HIP;
    int rc = PHYLO_OK;
    int *dl = nullptr, *dr = nullptr, *droot = nullptr;
    double *db = nullptr, *de = nullptr, *dp = nullptr, *dc = nullptr;
    float* df = nullptr;
    uint32_t* dm = nullptr;
    auto up = [&](auto*& d, const auto* h, size_t n) {
        if (hipMalloc(reinterpret_cast<void**>(&d), n * sizeof(*h)) != hipSuccess) return false;
        return hipMemcpyAsync(d, h, n * sizeof(*h), hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!up(dl, left, (size_t)n_align * nn) || !up(dr, right, (size_t)n_align * nn) || !up(droot, roots, (size_t)n_align) ||
        !up(db, blen, (size_t)n_align * nn) || !up(de, eig, (size_t)n_align * 36) || !up(dp, pi, (size_t)n_align * 4) ||
        !up(dc, cat_rates, (size_t)n_align * n_cat) ||
        hipMalloc(reinterpret_cast<void**>(&df), (size_t)n_align * kFeat * sizeof(float)) != hipSuccess ||
        (masks && hipMalloc(reinterpret_cast<void**>(&dm), (size_t)n_align * n_taxa * n_sites * sizeof(uint32_t)) !=
                      hipSuccess)) {
        rc = PHYLO_ERR_HIP;
    } else {
        hipLaunchKernelGGL(sim_features_kernel, dim3(n_align), dim3(kSimThreads), 0, st, n_taxa, n_sites, n_cat, dl, dr,
                           droot, db, de, dp, dc, (float)gap_frac, seed, df, dm);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(features, df, (size_t)n_align * kFeat * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
            (masks && hipMemcpyAsync(masks, dm, (size_t)n_align * n_taxa * n_sites * sizeof(uint32_t),
                                     hipMemcpyDeviceToHost, st) != hipSuccess) ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = PHYLO_ERR_HIP;
    }
    void* bufs[] = {dl, dr, droot, db, de, dp, dc, df, dm};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    (void)hipStreamDestroy(st);
    return rc;
}

int phylo_mlp_create(int device, int n_in, int n_hidden, int n_out, int max_batch, uint64_t seed, phylo_mlp** out) {
    if (!out || n_in < 1 || n_hidden < 1 || n_out < 1 || max_batch < 1) return PHYLO_ERR_ARG;
    *out = nullptr;
    phylo_mlp* p = new phylo_mlp();
    p->device = device;
    p->n_in = n_in;
    p->n_hid = n_hidden;
    p->n_out = n_out;
    p->max_batch = max_batch;
    p->n_params = (long)n_in * n_hidden + n_hidden + (long)n_hidden * n_hidden + n_hidden + (long)n_hidden * n_out + n_out;
    auto alloc = [&](auto*& d, size_t n) { return hipMalloc(reinterpret_cast<void**>(&d), n * sizeof(*d)) == hipSuccess; };
    const size_t B = max_batch;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        !alloc(p->d_w, p->n_params) || !alloc(p->d_g, p->n_params) || !alloc(p->d_m, p->n_params) ||
        !alloc(p->d_v, p->n_params) || !alloc(p->d_x, B * n_in) || !alloc(p->d_t, B * n_out) ||
        !alloc(p->d_h1, B * n_hidden) || !alloc(p->d_h2, B * n_hidden) || !alloc(p->d_y, B * n_out) ||
        !alloc(p->d_dh1, B * n_hidden) || !alloc(p->d_dh2, B * n_hidden) || !alloc(p->d_dy, B * n_out) ||
        !alloc(p->d_loss, nblk((long)B * n_out)) || !alloc(p->d_idx, B)) {
        phylo_mlp_destroy(p);
        return PHYLO_ERR_HIP;
    }
    // He-normal weights, zero biases
    std::vector<float> w(p->n_params, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> N01(0.f, 1.f);
    Views hv;
    hv.W1 = w.data();
    hv.b1 = hv.W1 + (size_t)n_in * n_hidden;
    hv.W2 = hv.b1 + n_hidden;
    hv.b2 = hv.W2 + (size_t)n_hidden * n_hidden;
    hv.W3 = hv.b2 + n_hidden;
    hv.b3 = hv.W3 + (size_t)n_hidden * n_out;
    for (long i = 0; i < (long)n_in * n_hidden; ++i) hv.W1[i] = N01(rng) * std::sqrt(2.f / n_in);
    for (long i = 0; i < (long)n_hidden * n_hidden; ++i) hv.W2[i] = N01(rng) * std::sqrt(2.f / n_hidden);
    for (long i = 0; i < (long)n_hidden * n_out; ++i) hv.W3[i] = N01(rng) * std::sqrt(1.f / n_hidden);
    std::vector<float> zero(p->n_params, 0.f);
    if (hipMemcpy(p->d_w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_m, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_v, zero.data(), zero.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        phylo_mlp_destroy(p);
        return PHYLO_ERR_HIP;
    }
    *out = p;
    return PHYLO_OK;
}

const char* phylo_mlp_last_error(const phylo_mlp* p) { return p ? p->err.c_str() : "null mlp"; }

long phylo_mlp_n_params(const phylo_mlp* p) { return p ? p->n_params : 0; }

int phylo_mlp_params(phylo_mlp* p, float* get, const float* set) {
    if (!p) return PHYLO_ERR_ARG;
    MCHK(p, hipSetDevice(p->device));
    if (set) MCHK(p, hipMemcpy(p->d_w, set, p->n_params * sizeof(float), hipMemcpyHostToDevice));
    if (get) MCHK(p, hipMemcpy(get, p->d_w, p->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return PHYLO_OK;
}

int phylo_mlp_forward(phylo_mlp* p, int batch, const float* x, float* y) {
    if (!p || !x || !y || batch < 1) return PHYLO_ERR_ARG;
    MCHK(p, hipSetDevice(p->device));
    for (int b0 = 0; b0 < batch; b0 += p->max_batch) {
        const int B = std::min(p->max_batch, batch - b0);
        MCHK(p, hipMemcpyAsync(p->d_x, x + (size_t)b0 * p->n_in, (size_t)B * p->n_in * sizeof(float),
                               hipMemcpyHostToDevice, p->stream));
        forward(p, B, p->d_x);
        MCHK(p, hipGetLastError());
        MCHK(p, hipMemcpyAsync(y + (size_t)b0 * p->n_out, p->d_y, (size_t)B * p->n_out * sizeof(float),
                               hipMemcpyDeviceToHost, p->stream));
    }
    MCHK(p, hipStreamSynchronize(p->stream));
    return PHYLO_OK;
}

int phylo_mlp_grad(phylo_mlp* p, int batch, const float* x, const float* t, float* loss, float* grad) {
    if (!p || !x || !t || batch < 1 || batch > p->max_batch) return PHYLO_ERR_ARG;
    MCHK(p, hipSetDevice(p->device));
    MCHK(p, hipMemcpyAsync(p->d_x, x, (size_t)batch * p->n_in * sizeof(float), hipMemcpyHostToDevice, p->stream));
    MCHK(p, hipMemcpyAsync(p->d_t, t, (size_t)batch * p->n_out * sizeof(float), hipMemcpyHostToDevice, p->stream));
    forward(p, batch, p->d_x);
    backward(p, batch, p->d_x, p->d_t);
    MCHK(p, hipGetLastError());
    if (grad) MCHK(p, hipMemcpyAsync(grad, p->d_g, p->n_params * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    MCHK(p, hipStreamSynchronize(p->stream));
    if (loss) *loss = read_loss(p, batch);
    return PHYLO_OK;
}

int phylo_mlp_set_data(phylo_mlp* p, int n, const float* x, const float* t) {
    if (!p || n < 1 || !x || !t) return PHYLO_ERR_ARG;
    MCHK(p, hipSetDevice(p->device));
    if (p->d_data_x) (void)hipFree(p->d_data_x);
    if (p->d_data_t) (void)hipFree(p->d_data_t);
    p->d_data_x = nullptr;
    p->d_data_t = nullptr;
    MCHK(p, hipMalloc(reinterpret_cast<void**>(&p->d_data_x), (size_t)n * p->n_in * sizeof(float)));
    MCHK(p, hipMalloc(reinterpret_cast<void**>(&p->d_data_t), (size_t)n * p->n_out * sizeof(float)));
    MCHK(p, hipMemcpy(p->d_data_x, x, (size_t)n * p->n_in * sizeof(float), hipMemcpyHostToDevice));
    MCHK(p, hipMemcpy(p->d_data_t, t, (size_t)n * p->n_out * sizeof(float), hipMemcpyHostToDevice));
    p->n_data = n;
    return PHYLO_OK;
}

// Epochs of minibatch Adam over the stored training set (shuffled per epoch).
int phylo_mlp_train(phylo_mlp* p, int epochs, int batch, float lr, uint64_t seed, float* loss_per_epoch) {
    if (!p || epochs < 1 || batch < 1 || batch > p->max_batch || !p->d_data_x) return PHYLO_ERR_ARG;
    MCHK(p, hipSetDevice(p->device));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(p->n_data);
    const float b1 = 0.9f, b2 = 0.999f;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int i = 0; i < p->n_data; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        double lsum = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= p->n_data; b0 += batch) {
            MCHK(p, hipMemcpyAsync(p->d_idx, perm.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, p->stream));
            hipLaunchKernelGGL(gather_rows_kernel, dim3(nblk((long)batch * p->n_in)), dim3(256), 0, p->stream, batch,
                               p->n_in, p->d_idx, p->d_data_x, p->d_x);
            hipLaunchKernelGGL(gather_rows_kernel, dim3(nblk((long)batch * p->n_out)), dim3(256), 0, p->stream, batch,
                               p->n_out, p->d_idx, p->d_data_t, p->d_t);
            forward(p, batch, p->d_x);
            backward(p, batch, p->d_x, p->d_t);
            ++p->step;
            const float c1 = 1.f - std::pow(b1, (float)p->step), c2 = 1.f - std::pow(b2, (float)p->step);
            hipLaunchKernelGGL(adam_kernel, dim3(nblk(p->n_params)), dim3(256), 0, p->stream, p->n_params, p->d_w,
                               p->d_g, p->d_m, p->d_v, lr, b1, b2, 1e-8f, c1, c2);
            MCHK(p, hipGetLastError());
            lsum += read_loss(p, batch);   // synchronises the stream (perm buffer reuse is safe)
            ++nb;
        }
        if (loss_per_epoch) loss_per_epoch[ep] = nb ? (float)(lsum / nb) : NAN;
    }
    return PHYLO_OK;
}

void phylo_mlp_destroy(phylo_mlp* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    void* bufs[] = {p->d_w, p->d_g, p->d_m, p->d_v, p->d_x, p->d_t, p->d_h1, p->d_h2, p->d_y,
                    p->d_dh1, p->d_dh2, p->d_dy, p->d_loss, p->d_idx, p->d_data_x, p->d_data_t};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

}  // extern "C"

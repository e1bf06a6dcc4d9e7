// Level-0 operator entry point that allocates no device memory.
//
// gputo_matvec_f64_noalloc is gputo_matvec_f64 (y = K(CX) u, y[fixed] = 0:
// multAV0c + ZeroBCc, Cantilever/TOP3D125MGCGCUDA/mgcg1.cu:941-997,1326)
// for callers that cannot call hipMalloc: code under HIP graph capture,
// serving loops.  The caller passes device inputs and outputs; the only
// device memory beyond them, the one-byte-per-node fixed-DOF mask the
// node-gather kernel reads, lives in `scratch`, a static device array of
// kScratchBytes per GPU, held until the process ends.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU race.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "gputo_internal.h"
#include "ke_classes.h"

namespace gputo {
namespace {

constexpr size_t kScratchBytes = 8u << 20;  // node masks of meshes up to 8,388,608 nodes
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
unsigned char *g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU
std::mutex g_scratch_mu;

// Address of this GPU's `scratch` (looked up once per device, with the
// device current); null on a HIP error.
unsigned char *scratch_ptr(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch[dev];
}

// mask words = 0 (a fill kernel rather than a memset node, so the call can
// be captured into a graph)
__global__ void k_mask_clear(uint32_t *m, long nwords)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long)gridDim.x * blockDim.x)
        m[i] = 0u;
}

// bit d of byte n set for each fixed DOF 3n+d; DOFs outside [0, ndof) are
// skipped.  Several DOFs of one node may be fixed, so the byte is set by an
// integer OR on its aligned word (order-independent).
__global__ void k_mask_set(uint8_t *mask, const int *fixed, int nfd, long ndof)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfd) return;
    const long d = fixed[i];
    if (d < 0 || d >= ndof) return;
    const long n = d / 3;
    const unsigned bit = 1u << (8 * (unsigned)(n & 3) + (unsigned)(d - 3 * n));
    atomicOr(reinterpret_cast<unsigned *>(mask + (n & ~3L)), bit);
}

}  // namespace
}  // namespace gputo

extern "C" int gputo_matvec_f64_noalloc(int nelx, int nely, int nelz, const double *Ke0, const double *d_CX,
                                        const double *d_u, const int *d_fixed, int nfd, double *d_y, void *stream)
{
    using namespace gputo;
    if (nelx < 1 || nely < 1 || nelz < 1 || !Ke0 || !d_CX || !d_u || !d_y || nfd < 0 || (nfd > 0 && !d_fixed))
        return GPUTO_ERR_ARGUMENT;
    const Grid g = make_grid(nelx, nely, nelz);
    const long nwords = (g.nnod + 3) / 4;
    if ((size_t)nwords * 4 > kScratchBytes) {
        std::fprintf(stderr, "gputo_matvec_f64_noalloc: %ld nodes exceed the %zu-node scratch mask\n", g.nnod,
                     kScratchBytes);
        return GPUTO_ERR_WORKSPACE;
    }
    // the node-gather kernel takes Ke0 as its entry classes (kernel
    // arguments), so no device copy of Ke0 is needed
    OpArgsT<double> a{};
    for (int k = 0; k < kKeClasses; ++k) a.kv[k] = Ke0[kKeFirst[k]];
    for (int q = 0; q < 576; ++q)
        if (std::memcmp(&Ke0[q], &a.kv[kKeClass[q]], sizeof(double)) != 0) {
            std::fprintf(stderr, "gputo_matvec_f64_noalloc: Ke0 lacks the unit-H8 entry structure\n");
            return GPUTO_ERR_ARGUMENT;
        }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return GPUTO_ERR_DEVICE;
    uint8_t *mask = scratch_ptr(dev);
    if (!mask) return GPUTO_ERR_DEVICE;
    hipStream_t s = static_cast<hipStream_t>(stream);
    k_mask_clear<<<(unsigned)std::min<long>((nwords + 255) / 256, 1024), 256, 0, s>>>(
        reinterpret_cast<uint32_t *>(mask), nwords);
    if (nfd > 0) k_mask_set<<<(nfd + 255) / 256, 256, 0, s>>>(mask, d_fixed, nfd, g.ndof);
    a.g = g;
    a.t = choose_tile(g);
    a.kc = 1;
    a.CX = d_CX;
    a.U = d_u;
    a.mask = mask;
    a.out = d_y;
    launch_op0<double>(s, OP_MATVEC, a);
    return hipGetLastError() == hipSuccess ? 0 : GPUTO_ERR_DEVICE;
}

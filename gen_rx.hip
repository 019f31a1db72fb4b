typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef float TT;
typedef float SS;
typedef cm::OpSum OP;
struct JitRxArgs { const void* leaf[12]; void* partials; void* out; long long L, chunk; int nchunk, pad; };
#define EL(j0, k) ((j0) + ((k) >> 2) * 256 + ((k) & 3))
extern "C" __global__ __launch_bounds__(256) void cm_jit_reduce(const JitRxArgs A)
{
    typedef uint32_t R;
    __shared__ TT lds[4];
    const long long s = blockIdx.x / A.nchunk;
    const int c = blockIdx.x % A.nchunk;
    const long long beg = c * A.chunk;
    const long long end = min(A.L, beg + A.chunk);
    const R c2 = ((const uint32_t*)A.leaf[2])[0];
    TT acc = 0.0f;
    const long long t0 = (long long)(threadIdx.x >> 6) * 512 + (threadIdx.x & 63) * 4;
    for (long long p0 = beg; p0 < end; p0 += 256 * 8) {
        const long long j0 = p0 + t0;
        R v0[8];
        R v1[8];
        const long long e0 = s * A.L + j0;
        if (j0 + 259 < end && (e0 & 3) == 0) {
            { const u4v a = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[0] + e0)); const u4v b = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[0] + e0 + 256));
              v0[0] = a.x; v0[1] = a.y; v0[2] = a.z; v0[3] = a.w; v0[4] = b.x; v0[5] = b.y; v0[6] = b.z; v0[7] = b.w; }
            { const u4v a = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[1] + e0)); const u4v b = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[1] + e0 + 256));
              v1[0] = a.x; v1[1] = a.y; v1[2] = a.z; v1[3] = a.w; v1[4] = b.x; v1[5] = b.y; v1[6] = b.z; v1[7] = b.w; }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) v0[k] = ((const uint32_t*)A.leaf[0])[s * A.L + (EL(j0, k) < end ? EL(j0, k) : beg)];
#pragma unroll
            for (int k = 0; k < 8; ++k) v1[k] = ((const uint32_t*)A.leaf[1])[s * A.L + (EL(j0, k) < end ? EL(j0, k) : beg)];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            R acc_r = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;
            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;
#define acc acc_r
acc = v0[k];
t = v1[k];
acc = binary1<CM_F32, 2>(acc, acc, t);
t = c2;
acc = binary1<CM_F32, 0>(acc, acc, t);
#undef acc
            if (EL(j0, k) < end) acc = OP::apply(acc, (TT)getf(acc_r));
        }
    }
    acc = cm::block_reduce<OP, 256>(acc, lds);
    if (threadIdx.x == 0) {
        if (A.nchunk == 1) ((SS*)A.out)[s] = (SS)acc;
        else ((TT*)A.partials)[blockIdx.x] = acc;
    }
}

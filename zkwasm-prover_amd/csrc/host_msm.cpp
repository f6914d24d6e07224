// Host-side mirror of the reference's MSM driver (src/cuda/msm.rs:239-455):
// the same window policy, buffer set and call sequence over the C-ABI, so the
// drop-in `msm` / `batch_msm_collect` are exercised exactly as the prover's
// host code drives them.  Streams use the default priority.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "guard.h"
#include "zkwasm_prover_kernel.h"

namespace {

uint32_t ceil_log2(uint64_t v) {
    uint32_t k = 0;
    while ((1ull << k) < v) k++;
    return k;
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16); }
};

#define HCHECK(x)                             \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

}  // namespace

extern "C" {

// Window policy of src/cuda/msm.rs:274-283 for `count` MSMs of `len` points.
void zkp_msm_policy(int count, int len, int *window_bits, int *windows, int *workers) {
    const int bits = 254, threads = 64;
    int wb = (int)(ceil_log2((uint64_t)count) < 3 ? ceil_log2((uint64_t)count) : 3) + 10 +
             (int)(ceil_log2((uint64_t)len) > 22 ? ceil_log2((uint64_t)len) : 22) - 22;
    *window_bits = wb;
    *windows = (bits + wb - 1) / wb;
    int max_worker = 128 * 512;
    *workers = len < max_worker ? (len + threads - 1) / threads * threads : max_worker;
}

// batch_msm (src/cuda/msm.rs:239-367) with device-resident scalar sets:
// out_affine receives `count` affine points {x, y} (Montgomery Fp, 64 B each,
// identity = (0, 0)).
hipError_t zkp_batch_msm(void *out_affine, const void *points_dev, const void *const *scalars_dev, int count,
                         int len, int skip_zero) {
    return zk::guard([&]() -> hipError_t {
        if (count <= 0) return hipSuccess;
        if (len <= 0 || !out_affine || !points_dev || !scalars_dev) return hipErrorInvalidValue;
        int window_bits, windows, workers;
        zkp_msm_policy(count, len, &window_bits, &windows, &workers);
        const size_t bucket_size = (size_t)windows << window_bits;
        const size_t curve_buf_size = bucket_size + 2 * (size_t)workers;  // Jacobian points per MSM
        const size_t jac = 96;
        DevBuf curve_buf, sort_buf, idx[4];
        std::vector<DevBuf> acc(count);
        HCHECK(curve_buf.alloc(curve_buf_size * jac * count));
        const size_t single = (size_t)len * windows;
        HCHECK(sort_buf.alloc(single * 3 * 4));
        for (auto &b : idx) HCHECK(b.alloc(single * 4));
        for (auto &b : acc) HCHECK(b.alloc((size_t)workers * 2 * 4));

        hipStream_t streams[2];
        HCHECK(hipStreamCreate(&streams[0]));
        HCHECK(hipStreamCreate(&streams[1]));
        hipError_t err = hipSuccess;
        for (int i = 0; i < count && err == hipSuccess; i++) {
            hipStream_t s = streams[i & 1], other = streams[1 - (i & 1)];
            // the index/sort buffers are shared: wait for the other stream first
            err = hipStreamSynchronize(other);
            if (err != hipSuccess) break;
            err = msm((uint8_t *)curve_buf.p + (size_t)i * curve_buf_size * jac, points_dev, scalars_dev[i], idx[0].p,
                      idx[1].p, idx[2].p, idx[3].p, sort_buf.p, acc[i].p, len, windows, window_bits, 64, workers,
                      (int)(single * 3 * 4 > 0x7fffffff ? 0x7fffffff : single * 3 * 4), skip_zero, s);
        }
        hipError_t e0 = hipStreamSynchronize(streams[0]);
        hipError_t e1 = hipStreamSynchronize(streams[1]);
        (void)hipStreamDestroy(streams[0]);
        (void)hipStreamDestroy(streams[1]);
        HCHECK(err);
        HCHECK(e0);
        HCHECK(e1);

        // batch_msm_acc (src/cuda/msm.rs:369-455): device arrays of pointers
        std::vector<uint64_t> ptrs(count);
        for (int i = 0; i < count; i++) ptrs[i] = (uint64_t)((uint8_t *)curve_buf.p + (size_t)i * curve_buf_size * jac);
        DevBuf ptr_dev;
        HCHECK(ptr_dev.alloc(count * 8));
        HCHECK(hipMemcpy(ptr_dev.p, ptrs.data(), count * 8, hipMemcpyHostToDevice));
        HCHECK(batch_msm_collect(nullptr, nullptr, nullptr, nullptr, ptr_dev.p, workers, windows, window_bits, count,
                                 nullptr));
        HCHECK(hipDeviceSynchronize());
        std::vector<uint8_t> res(jac);
        for (int i = 0; i < count; i++) {
            HCHECK(hipMemcpy(res.data(), (void *)ptrs[i], jac, hipMemcpyDeviceToHost));
            uint8_t *o = (uint8_t *)out_affine + (size_t)i * 64;
            // result is normalised (z = 1) or the identity (z = 0)
            bool zero_z = true;
            for (int k = 64; k < 96; k++) zero_z &= res[k] == 0;
            if (zero_z)
                memset(o, 0, 64);
            else
                memcpy(o, res.data(), 64);
        }
        return hipSuccess;
    });
}

}  // extern "C"

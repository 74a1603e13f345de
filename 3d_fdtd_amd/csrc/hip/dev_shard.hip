// Halo planes between z-slab shards of one large grid.  A shard is an fdtd_dev grid whose first
// and last z planes are ghost planes: the kernels never update a grid's boundary planes, so
// they hold the neighbours' values.  Per step, after the E update the upper shard's first owned
// plane of (Ex, Ey) goes to the lower shard's top ghost plane, and after the H update the lower
// shard's last owned plane of (Hx, Hy) goes to the upper shard's bottom ghost plane; the CPML
// z slabs belong to the end shards.  On one GPU the copy is a device-to-device copy on the
// destination grid's stream, ordered after the source grid's queued work; across GPUs the same
// planes would travel by RCCL send/recv or peer copies over xGMI.
#include "dev_internal.h"

using namespace fdtd;

extern "C" {

int fdtd_dev_copy_planes(fdtd_dev *dst, int kd, fdtd_dev *src, int ks, unsigned comp_mask) {
    if (!dst || !src || dst->I != src->I || dst->J != src->J) {
        set_error("copy_planes: grids must share the x-y cross-section");
        return -2;
    }
    if (kd < 0 || kd >= dst->K || ks < 0 || ks >= src->K || comp_mask == 0 || comp_mask > 63) {
        set_error("copy_planes: plane index outside the grid or empty component mask");
        return -2;
    }
    fields_changed(dst);
    // the copy waits for the source's queued work; the source's later work waits for the copy
    FDTD_TRY(hipEventRecord(src->xev, src->stream));
    FDTD_TRY(hipStreamWaitEvent(dst->stream, src->xev, 0));
    const size_t bytes = (size_t)dst->P * sizeof(float);
    for (int m = 0; m < 6; m++)
        if (comp_mask & (1u << m))
            FDTD_TRY(hipMemcpyAsync(dst->F[m] + (size_t)kd * dst->P, src->F[m] + (size_t)ks * src->P, bytes,
                                    hipMemcpyDeviceToDevice, dst->stream));
    FDTD_TRY(hipEventRecord(dst->xev, dst->stream));
    FDTD_TRY(hipStreamWaitEvent(src->stream, dst->xev, 0));
    dst->alt_valid = false;  // a fused step's second buffer no longer matches the ghost plane
    return 0;
}

}  // extern "C"

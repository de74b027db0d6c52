// tsdf_sched.hpp -- host helpers of the reintegration scheduler: the pose <->
// dof conversions the reference keeps per frame and the selection of the
// frames to reintegrate.  Included once, at the end of tsdf_hip.hip; host
// code only (no device work).
//
//   MattoDof                  matrixutil.cpp:38-64    tsdf_mat_to_dof
//   getT                      matrixutil.h:20-63      tsdf_dof_to_mat
//   diff                      integrate.cpp:211-216   tsdf_reint_diff
//   updateReintegrationBuffer integrate.cpp:223-234   tsdf_reint_select
#pragma once

#include <algorithm>
#include <vector>

namespace {

// 4x4 row-major double product o = a b, each coefficient summed in k order
void mul4(const double* a, const double* b, double* o)
{
    double t[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double s = a[4 * r] * b[c];
            for (int k = 1; k < 4; ++k)
                s += a[4 * r + k] * b[4 * k + c];
            t[4 * r + c] = s;
        }
    std::memcpy(o, t, sizeof(t));
}

}  // namespace

extern "C" {

tsdf_status tsdf_mat_to_dof(const double T[16], double dof[6])
{
    if (!T || !dof)
        return fail(TSDF_BAD_ARG, "NULL matrix or dof");
    // (rx, ry, rz) of R = Rz Ry Rx, then the translation
    dof[0] = std::atan2(T[9], T[10]);
    dof[1] = -std::asin(T[8]);
    dof[2] = std::atan2(T[4], T[0]);
    dof[3] = T[3];
    dof[4] = T[7];
    dof[5] = T[11];
    return TSDF_OK;
}

tsdf_status tsdf_dof_to_mat(const double dof[6], double T[16])
{
    if (!dof || !T)
        return fail(TSDF_BAD_ARG, "NULL dof or matrix");
    const double cx = std::cos(dof[0]), sx = std::sin(dof[0]);
    const double cy = std::cos(dof[1]), sy = std::sin(dof[1]);
    const double cz = std::cos(dof[2]), sz = std::sin(dof[2]);
    const double tr[16] = {1, 0, 0, dof[3], 0, 1, 0, dof[4], 0, 0, 1, dof[5], 0, 0, 0, 1};
    const double rz[16] = {cz, -sz, 0, 0, sz, cz, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const double ry[16] = {cy, 0, sy, 0, 0, 1, 0, 0, -sy, 0, cy, 0, 0, 0, 0, 1};
    const double rx[16] = {1, 0, 0, 0, 0, cx, -sx, 0, 0, sx, cx, 0, 0, 0, 0, 1};
    double a[16];
    mul4(tr, rz, a);  // ((trans Rz) Ry) Rx, the evaluation order of getT
    mul4(a, ry, a);
    mul4(a, rx, T);
    return TSDF_OK;
}

double tsdf_reint_diff(const double world_dof[6], const double integrated_dof[6])
{
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double d = world_dof[k] - integrated_dof[k];
        s += d * d;
    }
    for (int k = 3; k < 6; ++k) {
        const double d = 2.0 * (world_dof[k] - integrated_dof[k]);
        s += d * d;
    }
    return std::sqrt(s);
}

tsdf_status tsdf_reint_select(int32_t n, const double* world_dof, const double* integrated_dof, double treint,
                              int32_t* order, int32_t* n_selected)
{
    if (n < 0 || !n_selected || (n > 0 && (!world_dof || !integrated_dof || !order)))
        return fail(TSDF_BAD_ARG, "bad frame count or NULL array");
    std::vector<double> d((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        d[(size_t)i] = tsdf_reint_diff(world_dof + 6 * (size_t)i, integrated_dof + 6 * (size_t)i);
        order[i] = i;
    }
    // list::sort(compFrames) is stable: equal differences keep list order
    std::stable_sort(order, order + n, [&](int32_t a, int32_t b) { return d[(size_t)a] > d[(size_t)b]; });
    int32_t m = 0;
    while (m < n && !(d[(size_t)order[m]] < treint))
        ++m;
    *n_selected = m;
    return TSDF_OK;
}

}  // extern "C"

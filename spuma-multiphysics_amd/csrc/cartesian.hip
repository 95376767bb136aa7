// SPDX-License-Identifier: GPL-3.0-or-later
// Structured-box fast path of the transport stage.
//
// When the caller's face list is an axis-aligned nx*ny*nz box (uniform face
// area and cell volume, internal faces plus boundary faces of any kind:
// extrapolated, slip wall, fixed state), one fused kernel does Cells + Faces
// + Rhs + Advance: the HLL coefficients of every face are rebuilt from the
// two adjacent states (a boundary face's right state as View builds it), so
// q, p and c_frozen are streamed once per stage instead of through per-face
// work records.  Each face is evaluated once and its flux added to both
// cells with opposite signs, and each cell adds its six contributions in its
// CSR (face-id) order, given by the slot-order table of its boundary type.
// The result is bitwise the generic path's (reactiveTransportKernels.h:
// 134-149, 343-356, 502-526, 551-560).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <vector>

#include "hip_common.h"
#include "transport_internal.h"
#include "cartesian_kernels.h"

namespace ramd {

CartesianInfo Transport::detectCartesian(const ReactiveTransportFace* f, const double* volumes,
                                         const std::vector<uint32_t>& row, const std::vector<int32_t>& incidence) {
    CartesianInfo info;
    const size_t nc = cfg.cells, nf = cfg.faces;
    if (nc < 8 || !cartesianSupported(ns, nv)) return info;
    for (size_t c = 1; c < nc; ++c)
        if (std::memcmp(&volumes[c], &volumes[0], sizeof(double))) return info;
    // axis and side of every face; exact unit normals only
    std::vector<int8_t> axis(nf), side(nf);
    long nx = 0, nxy = 0;
    for (size_t i = 0; i < nf; ++i) {
        if (f[i].kind < 0 || f[i].kind > 3 || (f[i].kind == 0) != (f[i].neighbour >= 0)) return info;
        if (std::memcmp(&f[i].area, &f[0].area, sizeof(double))) return info;
        int a = -1, s = 0;
        for (int d = 0; d < 3; ++d) {
            const double v = f[i].normal[d];
            if (v == 1.0 || v == -1.0) {
                if (a >= 0) return info;
                a = d;
                s = v > 0 ? 1 : -1;
            } else if (v != 0.0) {
                return info;
            }
        }
        if (a < 0) return info;
        axis[i] = int8_t(a);
        side[i] = int8_t(s);
        if (f[i].kind == 0) {
            if (s < 0) return info;
            const long d = long(f[i].neighbour) - long(f[i].owner);
            if (a == 0 && d != 1) return info;
            if (a == 1) { if (nx && d != nx) return info; nx = d; }
            if (a == 2) { if (nxy && d != nxy) return info; nxy = d; }
        }
    }
    if (nx < 2 || nxy < 2 * nx || nxy % nx || long(nc) % nxy) return info;
    const int NX = int(nx), NY = int(nxy / nx), NZ = int(long(nc) / nxy);
    if (NZ < 2) return info;
    // slot face ids per cell: 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z
    std::vector<int32_t> slot(6 * nc, -1);
    // boundary kinds by side plane: -1 extrapolated, -2 slip wall, j fixed state j
    std::vector<int32_t> kinds(size_t(boundaryPlanes(NX, NY, NZ)), -1);
    bool special = false;
    const long strideOf[3] = {1, nx, nxy};
    const int extent[3] = {NX, NY, NZ};
    // exact normal bits of each slot class: [slot][boundary]
    double normal[6][2][3];
    bool seen[6][2] = {};
    for (size_t i = 0; i < nf; ++i) {
        const int a = axis[i];
        const long o = f[i].owner;
        const int coord[3] = {int(o % nx), int((o / nx) % (nxy / nx)), int(o / nxy)};
        int s0, bnd = f[i].kind != 0;
        if (!bnd) {
            if (coord[a] >= extent[a] - 1) return info;  // no wrap-around
            s0 = 2 * a + 1;
            const long nb = o + strideOf[a];
            if (slot[6 * o + s0] >= 0 || slot[6 * nb + 2 * a] >= 0) return info;
            slot[6 * o + s0] = int32_t(i);
            slot[6 * nb + 2 * a] = int32_t(i);
        } else {
            if (side[i] < 0 && coord[a] != 0) return info;
            if (side[i] > 0 && coord[a] != extent[a] - 1) return info;
            s0 = 2 * a + (side[i] > 0 ? 1 : 0);
            if (slot[6 * o + s0] >= 0) return info;
            slot[6 * o + s0] = int32_t(i);
            if (f[i].kind != 2) {
                if (f[i].kind == 3 && (f[i].fixed < 0 || size_t(f[i].fixed) >= cfg.fixed || f[i].fixed > INT32_MAX))
                    return info;
                kinds[size_t(boundaryPlane(NX, NY, NZ, s0, coord[0], coord[1], coord[2]))] =
                    f[i].kind == 1 ? -2 : int32_t(f[i].fixed);
                special = true;
            }
        }
        // an internal face is the owner's + slot and the neighbour's - slot
        const int classes[2] = {s0, bnd ? s0 : 2 * a};
        for (int s1 : classes) {
            if (!seen[s1][bnd]) {
                seen[s1][bnd] = true;
                for (int d = 0; d < 3; ++d) normal[s1][bnd][d] = f[i].normal[d];
            } else if (std::memcmp(normal[s1][bnd], f[i].normal, sizeof(f[i].normal))) {
                return info;
            }
        }
    }
    bool typeSeen[64] = {};
    for (size_t c = 0; c < nc; ++c) {
        int ids[6];
        for (int s = 0; s < 6; ++s) {
            ids[s] = slot[6 * c + s];
            if (ids[s] < 0) return info;
        }
        if (row[c + 1] - row[c] != 6) return info;
        const int i = int(c % nx), j = int((c / nx) % NY), k = int(c / nxy);
        const int type = (i == 0) | (i == NX - 1) << 1 | (j == 0) << 2 | (j == NY - 1) << 3 | (k == 0) << 4 |
                         (k == NZ - 1) << 5;
        // CSR order = ascending face id; find each position's slot
        uint32_t code = 0;
        for (int p = 0; p < 6; ++p) {
            const int32_t entry = incidence[row[c] + p];
            const int32_t id = entry < 0 ? -entry - 1 : entry - 1;
            int s = 0;
            while (s < 6 && ids[s] != id) ++s;
            if (s == 6) return info;
            code |= uint32_t(s) << (3 * p);
        }
        if (!typeSeen[type]) {
            typeSeen[type] = true;
            info.slotOrder[type] = code;
        } else if (info.slotOrder[type] != code) {
            return info;
        }
    }
    for (int s = 0; s < 6; ++s)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 3; ++d) info.normal[s][b][d] = seen[s][b] ? normal[s][b][d] : 0.0;
    info.nx = NX;
    info.ny = NY;
    info.nz = NZ;
    info.area = f[0].area;
    info.inverseVolume = 1 / volumes[0];
    info.volume = volumes[0];
    info.valid = true;
    info.fixedBase = (long long)cfg.cells;
    if (special) {
        info.boundaryKind = allocate<int32_t>(kinds.size());
        RAMD_HIP(hipMemcpy(info.boundaryKind, kinds.data(), kinds.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    BoxDev box;
    box.a = makeCartesianArgs(info, cfg.waveFactor);
    box.offset = 0;
    box.cells = nc;
    box.volume = info.volume;
    info.stageBlocks = tileBlocks(box.a);
    info.cflBlocks = long((nc + 255) / 256);
    const long long starts[4] = {0, info.stageBlocks, 0, info.cflBlocks};
    info.box = allocate<BoxDev>(1);
    info.starts = allocate<long long>(4);
    RAMD_HIP(hipMemcpy(info.box, &box, sizeof(box), hipMemcpyHostToDevice));
    RAMD_HIP(hipMemcpy(info.starts, starts, sizeof(starts), hipMemcpyHostToDevice));
    pressure = allocate<double>(all);
    sound = allocate<double>(all);
    return info;
}

__global__ void extractStateKernel(const ReactiveTransportState* __restrict__ s, size_t n, double* __restrict__ p,
                                   double* __restrict__ c, double* __restrict__ r) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = s[i].p;
    c[i] = s[i].sound;
    if (r) r[i] = s[i].rho;
}

void Transport::cartesianAdvance(double dt, int stage) {
    if (scheme == kHllc && !density) density = allocate<double>(all);
    // p, c_frozen (and rho) of the cells and of the fixed states
    hipLaunchKernelGGL(extractStateKernel, dim3(unsigned((all + 255) / 256)), dim3(256), 0, stream, state, all,
                       pressure, sound, scheme == kHllc ? density : nullptr);
    RAMD_HIP(hipGetLastError());
    launched();
    launchBatchStage(ns, nv, scheme, cartesian.box, cartesian.starts, 1, cartesian.stageBlocks, q, all, pressure,
                     sound, density, dt, nullptr, stage, initial, fluxFailures, stream);
    launched();
}

}  // namespace ramd

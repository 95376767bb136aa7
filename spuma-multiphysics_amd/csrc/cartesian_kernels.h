// SPDX-License-Identifier: GPL-3.0-or-later
// Fused structured-box stage and CFL kernels (see cartesian.hip for the
// parity argument): Cells + Faces + Rhs + Advance of one SSPRK2 stage in one
// launch, each face's HLL (or HLLC) coefficients rebuilt from the two
// adjacent states and every cell's update written once.  Workgroups are
// remapped so that each of the eight XCDs sweeps one contiguous z-range: the
// +/-nx*ny neighbours a tile reads were then loaded by the same XCD moments
// earlier and hit its L2.
#ifndef REACTIVE_AMD_CARTESIAN_KERNELS_H
#define REACTIVE_AMD_CARTESIAN_KERNELS_H
#include <hip/hip_runtime.h>
#include <cstdint>
#include "transport_internal.h"
#include "flux_device.h"

namespace ramd {

// Index of the boundary face on side `slot` of cell (i, j, k) in the side planes:
// -x, +x (j + ny k), -y, +y (i + nx k), -z, +z (i + nx j), in slot order.
__host__ __device__ inline long boundaryPlane(int nx, int ny, int nz, int slot, int i, int j, int k) {
    const long yz = long(ny) * nz, xz = long(nx) * nz, xy = long(nx) * ny;
    switch (slot) {
        case 0: return j + long(ny) * k;
        case 1: return yz + j + long(ny) * k;
        case 2: return 2 * yz + i + long(nx) * k;
        case 3: return 2 * yz + xz + i + long(nx) * k;
        case 4: return 2 * (yz + xz) + i + long(nx) * j;
        default: return 2 * (yz + xz) + xy + i + long(nx) * j;
    }
}
inline long boundaryPlanes(int nx, int ny, int nz) {
    return 2 * (long(ny) * nz + long(nx) * nz + long(nx) * ny);
}
__device__ inline int boundaryKindOf(const CartesianArgs& a, int slot, int i, int j, int k) {
    return a.boundaryKind ? a.boundaryKind[boundaryPlane(a.nx, a.ny, a.nz, slot, i, j, k)] : -1;
}

inline CartesianArgs makeCartesianArgs(const CartesianInfo& c, double waveFactor) {
    CartesianArgs a;
    a.boundaryKind = c.boundaryKind;
    a.fixedBase = c.fixedBase;
    a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
    a.area = c.area; a.invV = c.inverseVolume; a.waveFactor = waveFactor;
    for (int t = 0; t < 64; ++t) a.slotOrder[t] = c.slotOrder[t];
    for (int s = 0; s < 6; ++s)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 3; ++d) a.normal[s][b][d] = c.normal[s][b][d];
    return a;
}

constexpr int kXcds = 8;

// Logical block of a physical workgroup: physical blocks are dispatched
// round-robin over the XCDs, so XCD x runs logical blocks [x*per, (x+1)*per).
__device__ inline long xcdBlock(long physical, long logicalBlocks) {
    const long per = (logicalBlocks + kXcds - 1) / kXcds;
    return (physical % kXcds) * per + physical / kXcds;
}
inline unsigned xcdGrid(long logicalBlocks) {
    const long per = (logicalBlocks + kXcds - 1) / kXcds;
    return unsigned(per * kXcds);
}

// ---- tiled stage: every face computed once -------------------------------
//
// A workgroup owns a kTileX x kTileY tile of (i, j) columns over kChunkZ
// cells in k and marches up in k.  The +x and +y faces are computed by the
// cell on their minus side and shared through LDS, the +z face by the
// column's own thread (it is the next cell's -z face); a tile's -x / -y edge
// cells compute their -x / -y faces themselves, from the neighbour tile's
// cells.  Velocities are computed once per cell.  Each face is evaluated
// with the generic path's arithmetic and every cell sums its six faces in
// its CSR order, so the result is the same bit for bit.
// Cells per column a workgroup marches (the chunk depth CZ): 2, 4 or 8,
// chosen per box by the step's autotuner; RAMD_TILE_Z is the default.
#ifndef RAMD_TILE_Z
#define RAMD_TILE_Z 4
#endif
constexpr int kTileX = 32, kTileY = 8, kChunkZ = RAMD_TILE_Z;

// 1: interior cells whose faces sum in kOwnerNeighbourOrder take a path with
// no per-value choice of face (the same sum, bit for bit).
#ifndef RAMD_STAGE_INTERIOR
#define RAMD_STAGE_INTERIOR 1
#endif
// Slot order (3 bits per CSR position, slots 0..5 = -x +x -y +y -z +z) of an
// interior cell when internal faces are numbered by owner, then neighbour
// (an upper-triangular face list): the faces it neighbours (-z, -y, -x, by
// their owners' index), then the faces it owns (+x, +y, +z).
constexpr uint32_t kOwnerNeighbourOrder = 4u | 2u << 3 | 0u << 6 | 1u << 9 | 3u << 12 | 5u << 15;
constexpr int kChunkDepths[3] = {2, 4, 8};

template <int NS, int NV, bool HLLC>
struct TileCell {
    // r: recovered density (HLLC only).  11 doubles at nv = 5: a 22-dword
    // stride between the cells of a row in LDS, which spreads them over more
    // banks than 24 dwords would (3.5 % faster stage at 160^3)
    double q[NV], p, c, r, u[3];
};

// Raw loads of cell c (q, p, c_frozen, r), then the derived rho and u.
template <int NS, int NV, bool HLLC>
__device__ inline void loadTileRaw(TileCell<NS, NV, HLLC>& t, const double* __restrict__ q, size_t stride,
                                   const double* __restrict__ pArr, const double* __restrict__ cArr,
                                   const double* __restrict__ rArr, long c) {
    for (int v = 0; v < NV; ++v) t.q[v] = q[size_t(v) * stride + c];
    t.p = pArr[c];
    t.c = cArr[c];
    t.r = HLLC ? rArr[c] : 0.0;
}
template <int NS, int NV, bool HLLC>
__device__ inline void deriveTileCell(TileCell<NS, NV, HLLC>& t) {
    double rho = 0;
    for (int s = 0; s < NS; ++s) rho += t.q[s];
    for (int d = 0; d < 3; ++d) t.u[d] = t.q[NS + d] / rho;
}
template <int NS, int NV, bool HLLC>
__device__ inline void loadTileCell(TileCell<NS, NV, HLLC>& t, const double* __restrict__ q, size_t stride,
                                    const double* __restrict__ pArr, const double* __restrict__ cArr,
                                    const double* __restrict__ rArr, long c) {
    loadTileRaw(t, q, stride, pArr, cArr, rArr, c);
    deriveTileCell(t);
}

// Right state of a boundary face of cell `cur` (View::rightCell/rightValue/
// velocities, reactiveTransportKernels.h:134-149): the cell itself
// (extrapolated), the cell with its normal velocity and momentum mirrored
// (slip wall, the generic path's operation order), or a fixed state.
template <int NS, int NV, bool HLLC>
__device__ inline void boundaryRight(const CartesianArgs& a, int kind, const double* n,
                                     const TileCell<NS, NV, HLLC>& cur, TileCell<NS, NV, HLLC>& R,
                                     const double* __restrict__ q, size_t stride, const double* __restrict__ pArr,
                                     const double* __restrict__ cArr, const double* __restrict__ rArr) {
    if (kind >= 0) {
        loadTileCell(R, q, stride, pArr, cArr, rArr, long(a.fixedBase) + kind);
        return;
    }
    R = cur;
    if (kind == -2) {
        double dot = 0, normal = 0;
        for (int d = 0; d < 3; ++d) {
            dot += R.u[d] * n[d];
            normal += cur.q[NS + d] * n[d];
        }
        for (int d = 0; d < 3; ++d) {
            R.u[d] -= 2 * dot * n[d];
            R.q[NS + d] -= 2 * normal * n[d];
        }
    }
}

// flux * area of one face, L = owner side, R = right side (velocities precomputed)
template <int NS, int NV, bool HLLC>
__device__ inline void tileFace(const CartesianArgs& a, const TileCell<NS, NV, HLLC>& L,
                                const TileCell<NS, NV, HLLC>& R, const double* n, double* F, uint32_t& rejected) {
    double unL = 0, unR = 0;
    for (int d = 0; d < 3; ++d) {
        unL += L.u[d] * n[d];
        unR += R.u[d] * n[d];
    }
    const double aL = a.waveFactor * L.c, aR = a.waveFactor * R.c;
    const double left = unL - aL < unR - aR ? unL - aL : unR - aR;
    const double right = unL + aL > unR + aR ? unL + aL : unR + aR;
    const double wl = 0 < left ? 0 : left, wr = 0 > right ? 0 : right;
    const double invWave = 1 / (wr - wl);
    double4 w;
    w.x = wr * (unL - wl) * invWave;
    w.y = wl * (wr - unR) * invWave;
    w.z = (wr * L.p - wl * R.p) * invWave;
    w.w = (wr * L.p * unL - wl * R.p * unR) * invWave;
    if (HLLC && !hllcCoefficients(L.r, L.u, L.p, aL, L.q[NS + 3], R.r, R.u, R.p, aR, R.q[NS + 3], n, w)) ++rejected;
    for (int v = 0; v < NV; ++v) {
        double flux = w.x * L.q[v] + w.y * R.q[v];
        if (v >= NS && v < NS + 3) flux += (w.z - 0.0) * n[v - NS];
        if (v == NS + 3) flux += w.w;
        if (v < NS) flux += 0.0;
        else if (v < NS + 3) flux -= 0.0;
        else if (v == NS + 3) flux += 0.0;
        F[v] = flux * a.area;
    }
}

template <int NS, int NV, bool HLLC>
struct TileShared {
    TileCell<NS, NV, HLLC> cell[kTileY][kTileX];
    double faceX[kTileY][kTileX][NV], faceY[kTileY][kTileX][NV];
};

inline __host__ __device__ long tileBlocks(const CartesianArgs& a, int cz = kChunkZ) {
    return long((a.nx + kTileX - 1) / kTileX) * ((a.ny + kTileY - 1) / kTileY) * ((a.nz + cz - 1) / cz);
}

// Logical tile block b (< tileBlocks(a)) of one box; every thread of the
// workgroup calls it (it synchronises the workgroup).
template <int NS, int NV, bool HLLC, int CZ = kChunkZ>
__device__ inline void tiledStageBlock(const CartesianArgs& a, long b, TileShared<NS, NV, HLLC>& lds,
                                       const double* __restrict__ q, size_t stride, const double* __restrict__ pArr,
                                       const double* __restrict__ cArr, const double* __restrict__ rArr, double dt,
                                       int mode, double* __restrict__ out, uint32_t& rejected) {
    using Cell = TileCell<NS, NV, HLLC>;
    auto& shared = lds.cell;
    auto& faceX = lds.faceX;
    auto& faceY = lds.faceY;
    const int tilesX = (a.nx + kTileX - 1) / kTileX, tilesY = (a.ny + kTileY - 1) / kTileY;
    const int bx = int(b % tilesX), by = int((b / tilesX) % tilesY), bz = int(b / (long(tilesX) * tilesY));
    const int tx = threadIdx.x % kTileX, ty = threadIdx.x / kTileX;
    const int i = bx * kTileX + tx, j = by * kTileY + ty;
    const bool active = i < a.nx && j < a.ny;
    const int k0 = bz * CZ, k1 = min(a.nz, k0 + CZ);
    const long nxy = long(a.nx) * a.ny;
    Cell cur, nxt;
    double fzMinus[NV];
    if (active) {
        const long c0 = long(i) + long(a.nx) * j + nxy * k0;
        loadTileCell(cur, q, stride, pArr, cArr, rArr, c0);
        if (k0 == 0) {
            Cell ghost;
            boundaryRight(a, boundaryKindOf(a, 4, i, j, 0), a.normal[4][1], cur, ghost, q, stride, pArr, cArr, rArr);
            tileFace(a, cur, ghost, a.normal[4][1], fzMinus, rejected);
        } else {
            Cell below;
            loadTileCell(below, q, stride, pArr, cArr, rArr, c0 - nxy);
            tileFace(a, below, cur, a.normal[5][0], fzMinus, rejected);
        }
    }
    for (int k = k0; k < k1; ++k) {
        const long c = long(i) + long(a.nx) * j + nxy * k;
        double base[NV];  // stage 1: q0 of this cell, loaded before the face work so its latency overlaps it
        if (active && mode != 0)
            for (int v = 0; v < NV; ++v) base[v] = out[size_t(v) * stride + c];
        if (active) shared[ty][tx] = cur;
        __syncthreads();
        double fxMinus[NV], fyMinus[NV], fzPlus[NV];
        if (active) {
            double F[NV];
            // +x face (this cell is its owner)
            if (i == a.nx - 1) {
                Cell ghost;
                boundaryRight(a, boundaryKindOf(a, 1, i, j, k), a.normal[1][1], cur, ghost, q, stride, pArr, cArr,
                              rArr);
                tileFace(a, cur, ghost, a.normal[1][1], F, rejected);
            } else if (tx + 1 < kTileX) {
                tileFace(a, cur, shared[ty][tx + 1], a.normal[1][0], F, rejected);
            } else {
                Cell nb;
                loadTileCell(nb, q, stride, pArr, cArr, rArr, c + 1);
                tileFace(a, cur, nb, a.normal[1][0], F, rejected);
            }
            for (int v = 0; v < NV; ++v) faceX[ty][tx][v] = F[v];
            // +y face
            if (j == a.ny - 1) {
                Cell ghost;
                boundaryRight(a, boundaryKindOf(a, 3, i, j, k), a.normal[3][1], cur, ghost, q, stride, pArr, cArr,
                              rArr);
                tileFace(a, cur, ghost, a.normal[3][1], F, rejected);
            } else if (ty + 1 < kTileY) {
                tileFace(a, cur, shared[ty + 1][tx], a.normal[3][0], F, rejected);
            } else {
                Cell nb;
                loadTileCell(nb, q, stride, pArr, cArr, rArr, c + a.nx);
                tileFace(a, cur, nb, a.normal[3][0], F, rejected);
            }
            for (int v = 0; v < NV; ++v) faceY[ty][tx][v] = F[v];
            // +z face; the next cell of the column stays in registers
            if (k == a.nz - 1) {
                Cell ghost;
                boundaryRight(a, boundaryKindOf(a, 5, i, j, k), a.normal[5][1], cur, ghost, q, stride, pArr, cArr,
                              rArr);
                tileFace(a, cur, ghost, a.normal[5][1], fzPlus, rejected);
            } else {
                loadTileCell(nxt, q, stride, pArr, cArr, rArr, c + nxy);
                tileFace(a, cur, nxt, a.normal[5][0], fzPlus, rejected);
            }
            // -x / -y faces: domain boundary, or a tile edge (neighbour in another tile)
            if (i == 0) {
                Cell ghost;
                boundaryRight(a, boundaryKindOf(a, 0, i, j, k), a.normal[0][1], cur, ghost, q, stride, pArr, cArr,
                              rArr);
                tileFace(a, cur, ghost, a.normal[0][1], fxMinus, rejected);
            } else if (tx == 0) {
                Cell nb;
                loadTileCell(nb, q, stride, pArr, cArr, rArr, c - 1);
                tileFace(a, nb, cur, a.normal[1][0], fxMinus, rejected);
            }
            if (j == 0) {
                Cell ghost;
                boundaryRight(a, boundaryKindOf(a, 2, i, j, k), a.normal[2][1], cur, ghost, q, stride, pArr, cArr,
                              rArr);
                tileFace(a, cur, ghost, a.normal[2][1], fyMinus, rejected);
            } else if (ty == 0) {
                Cell nb;
                loadTileCell(nb, q, stride, pArr, cArr, rArr, c - a.nx);
                tileFace(a, nb, cur, a.normal[3][0], fyMinus, rejected);
            }
        }
        __syncthreads();
        if (active) {
            if (i > 0 && tx > 0)
                for (int v = 0; v < NV; ++v) fxMinus[v] = faceX[ty][tx - 1][v];
            if (j > 0 && ty > 0)
                for (int v = 0; v < NV; ++v) fyMinus[v] = faceY[ty - 1][tx][v];
            const int type = (i == 0) | (i == a.nx - 1) << 1 | (j == 0) << 2 | (j == a.ny - 1) << 3 | (k == 0) << 4 |
                             (k == a.nz - 1) << 5;
            const uint32_t order = a.slotOrder[type];
            double rate[NV];
            for (int v = 0; v < NV; ++v) rate[v] = 0;
            if (RAMD_STAGE_INTERIOR && type == 0 && order == kOwnerNeighbourOrder) {
                // interior cell of an owner-then-neighbour face list: the
                // generic sum below in its order, with no per-value choice of face
                for (int v = 0; v < NV; ++v) {
                    rate[v] += 1.0 * fzMinus[v] * a.invV;
                    rate[v] += 1.0 * fyMinus[v] * a.invV;
                    rate[v] += 1.0 * fxMinus[v] * a.invV;
                    rate[v] += -1.0 * faceX[ty][tx][v] * a.invV;
                    rate[v] += -1.0 * faceY[ty][tx][v] * a.invV;
                    rate[v] += -1.0 * fzPlus[v] * a.invV;
                }
            } else
            for (int pos = 0; pos < 6; ++pos) {
                const int slot = int((order >> (3 * pos)) & 7u);
                const int ax = slot >> 1, plus = slot & 1;
                const int coord = ax == 0 ? i : (ax == 1 ? j : k);
                const int ext = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
                const bool bnd = plus ? coord == ext - 1 : coord == 0;
                const double sign = bnd || plus ? -1.0 : 1.0;
                for (int v = 0; v < NV; ++v) {
                    const double F = slot == 0 ? fxMinus[v] : slot == 1 ? faceX[ty][tx][v] : slot == 2 ? fyMinus[v]
                                     : slot == 3 ? faceY[ty][tx][v] : slot == 4 ? fzMinus[v] : fzPlus[v];
                    rate[v] += sign * F * a.invV;
                }
            }
            for (int v = 0; v < NV; ++v) {
                const size_t idx = size_t(v) * stride + c;
                if (mode == 0) out[idx] = cur.q[v] + dt * rate[v];
                else out[idx] = .5 * base[v] + .5 * (cur.q[v] + dt * rate[v]);
            }
            for (int v = 0; v < NV; ++v) fzMinus[v] = fzPlus[v];
            cur = nxt;
        }
        __syncthreads();
    }
}



// FaceSpeeds + Step + the block level of Minimum (reactiveTransportKernels.h:
// 315-342, 529-550; reactiveTransport.cpp:222-232) on the structured box.
// Face speeds are rebuilt per cell and summed in CSR order; invalid cell
// steps map to -1 (absorbing).  maximum comes from maxDev when non-null.
template <int NS>
__device__ inline double cartesianCellStep(const CartesianArgs& a, const double* __restrict__ q, size_t stride,
                                           const double* __restrict__ cArr, long c, double cfl, double maximum) {
    double v;
    {
        const long nxy = long(a.nx) * a.ny;
        const int i = int(c % a.nx), j = int((c / a.nx) % a.ny), k = int(c / nxy);
        const int type = (i == 0) | (i == a.nx - 1) << 1 | (j == 0) << 2 | (j == a.ny - 1) << 3 | (k == 0) << 4 |
                         (k == a.nz - 1) << 5;
        const uint32_t order = a.slotOrder[type];
        double qc[NS + 3];
        for (int s = 0; s < NS + 3; ++s) qc[s] = q[size_t(s) * stride + c];
        const double cc = cArr[c];
        double denominator = 0, step;
        bool bad = false;
        for (int pos = 0; pos < 6 && !bad; ++pos) {
            const int slot = int((order >> (3 * pos)) & 7u);
            const int ax = slot >> 1, plus = slot & 1;
            const int coord = ax == 0 ? i : (ax == 1 ? j : k);
            const int ext = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
            const bool bnd = plus ? coord == ext - 1 : coord == 0;
            const long stepIdx = ax == 0 ? 1 : (ax == 1 ? long(a.nx) : nxy);
            const double* n = a.normal[slot][bnd ? 1 : 0];
            double qn[NS + 3], cn = cc;
            const int kind = bnd ? boundaryKindOf(a, slot, i, j, k) : -1;
            if (bnd && kind < 0) {
                for (int s = 0; s < NS + 3; ++s) qn[s] = qc[s];
            } else {
                const long nb = bnd ? long(a.fixedBase) + kind : (plus ? c + stepIdx : c - stepIdx);
                for (int s = 0; s < NS + 3; ++s) qn[s] = q[size_t(s) * stride + nb];
                cn = cArr[nb];
            }
            const bool selfLeft = bnd || plus;
            double rhoL = 0, rhoR = 0, unL = 0, unR = 0;
            for (int s = 0; s < NS; ++s) {
                rhoL += selfLeft ? qc[s] : qn[s];
                rhoR += selfLeft ? qn[s] : qc[s];
            }
            if (kind == -2) {  // slip wall: mirrored right velocity (generic operation order)
                double ur[3], dot = 0;
                for (int d = 0; d < 3; ++d) {
                    ur[d] = qn[NS + d] / rhoR;
                    dot += ur[d] * n[d];
                }
                for (int d = 0; d < 3; ++d) ur[d] -= 2 * dot * n[d];
                for (int d = 0; d < 3; ++d) {
                    unL += qc[NS + d] / rhoL * n[d];
                    unR += ur[d] * n[d];
                }
            } else {
                for (int d = 0; d < 3; ++d) {
                    unL += (selfLeft ? qc[NS + d] : qn[NS + d]) / rhoL * n[d];
                    unR += (selfLeft ? qn[NS + d] : qc[NS + d]) / rhoR * n[d];
                }
            }
            const double aL = a.waveFactor * (selfLeft ? cc : cn), aR = a.waveFactor * (selfLeft ? cn : cc);
            const double left0 = unL - aL < unR - aR ? unL - aL : unR - aR;
            const double right0 = unL + aL > unR + aR ? unL + aL : unR + aR;
            const double left = 0 < left0 ? 0 : left0, right = 0 > right0 ? 0 : right0;
            const double invWave = 1 / (right - left);
            const double fl = ::fabs(unL) + aL, fr = ::fabs(unR) + aR;
            double speed = (fl > fr ? fl : fr) + 0.0;
            if (!isfinite(unL) || !isfinite(unR) || !isfinite(speed) || !isfinite(invWave) || right <= left) speed = -1;
            if (!isfinite(speed) || speed < 0) { bad = true; break; }
            denominator += a.area * speed;
        }
        if (bad) step = -1;
        else {
            const double r = cfl / (a.invV * denominator);
            step = denominator > 0 ? (maximum < r ? maximum : r) : maximum;
            if (!isfinite(denominator) || denominator < 0) step = -1;
        }
        v = isfinite(step) && step > 0 ? step : -1;
    }
    return v;
}

// Minimum of v over a 256-thread workgroup, written to *out by thread 0.
__device__ inline void workgroupMinTo(double v, double* __restrict__ out) {
    __shared__ double lane[4];
    for (int offset = 32; offset > 0; offset >>= 1) {
        const double o = __shfl_xor(v, offset, 64);
        v = o < v ? o : v;
    }
    if ((threadIdx.x & 63) == 0) lane[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = lane[0];
        for (int w = 1; w < 4; ++w) m = lane[w] < m ? lane[w] : m;
        *out = m;
    }
}

inline bool cartesianSupported(int ns, int nv) {
    return (ns == 1 && (nv == 5 || nv == 6)) || (ns == 2 && (nv == 6 || nv == 7));
}

// ---- launches over a table of boxes ---------------------------------------
//
// Every launch sweeps a device table of boxes (BoxDev): box i owns cells
// [offset, offset + nx ny nz) of every SoA array (one common stride) and
// logical blocks [start[i], start[i + 1]) of the launch; a single box is a
// table of one.  Each box is swept with the same arithmetic whatever the
// table holds, so a box's results do not depend on the other boxes, bit for
// bit.  The box descriptors stay in device memory (a by-value kernel argument
// indexed per thread would be copied to scratch), and physical workgroups are
// remapped so each XCD sweeps one contiguous range of logical blocks.
// Box of logical block b: the last i with start[i] <= b (uniform per workgroup).
__device__ inline int boxOf(const long long* __restrict__ start, int boxes, long b) {
    int lo = 0, hi = boxes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// CFL cell steps of every box into blockMin[global block]; the maximum step
// of box i is maxDev[4 i] when maxDev is given.
template <int NS>
__global__ __launch_bounds__(256) void batchCflKernel(const BoxDev* __restrict__ boxes,
                                                      const long long* __restrict__ start, int count,
                                                      const double* __restrict__ q, size_t stride,
                                                      const double* __restrict__ cArr, double cfl, double maxHost,
                                                      const double* __restrict__ maxDev,
                                                      double* __restrict__ blockMin) {
    const long total = long(start[count]);
    const long b = xcdBlock(blockIdx.x, total);
    if (b >= total) return;  // uniform per workgroup
    const int i = boxOf(start, count, b);
    const BoxDev& box = boxes[i];
    const long c = (b - start[i]) * 256 + threadIdx.x;
    double v = __builtin_huge_val();
    if (c < long(box.cells))
        v = cartesianCellStep<NS>(box.a, q + box.offset, stride, cArr + box.offset, c, cfl,
                                  maxDev ? maxDev[4 * i] : maxHost);
    workgroupMinTo(v, blockMin + b);
}

// One stage of every box; box i advances with dt = dtDev[4 i] (dtHost when
// dtDev is null; a box whose dt is invalid is left untouched).  Three
// workgroups per CU fit the LDS at nv = 5 and the registers are held to
// three waves per SIMD for them; wider states run two.
template <int NS, int NV, bool HLLC, int CZ>
__global__ __launch_bounds__(kTileX * kTileY, NV <= 5 ? 3 : 2) void batchTiledStageKernel(
    const BoxDev* __restrict__ boxes, const long long* __restrict__ start, int count, const double* __restrict__ q,
    size_t stride, const double* __restrict__ pArr, const double* __restrict__ cArr, const double* __restrict__ rArr,
    double dtHost, const double* __restrict__ dtDev, int mode, double* __restrict__ out,
    unsigned long long* __restrict__ failures) {
    __shared__ TileShared<NS, NV, HLLC> lds;
    const long total = long(start[count]);
    const long b = xcdBlock(blockIdx.x, total);
    if (b >= total) return;  // uniform per workgroup
    const int i = boxOf(start, count, b);
    const BoxDev& box = boxes[i];
    const double dt = dtDev ? dtDev[4 * i] : dtHost;
    if (!(dt > 0)) return;  // uniform per workgroup
    const size_t o = box.offset;
    uint32_t rejected = 0;
    tiledStageBlock<NS, NV, HLLC, CZ>(box.a, b - start[i], lds, q + o, stride, pArr + o, cArr + o,
                                  HLLC ? rArr + o : nullptr, dt, mode, out + o, rejected);
    if (HLLC && rejected) atomicAdd(failures, (unsigned long long)rejected);
}

// CFL block minima of every box (blocks = start[count] of the CFL table).
inline void launchBatchCfl(int ns, const BoxDev* boxes, const long long* start, int count, long blocks,
                           const double* q, size_t stride, const double* c, double cfl, double maxHost,
                           const double* maxDev, double* blockMin, hipStream_t stream) {
    const unsigned grid = xcdGrid(blocks);
    if (ns == 1)
        hipLaunchKernelGGL((batchCflKernel<1>), dim3(grid), dim3(256), 0, stream, boxes, start, count, q, stride, c,
                           cfl, maxHost, maxDev, blockMin);
    else
        hipLaunchKernelGGL((batchCflKernel<2>), dim3(grid), dim3(256), 0, stream, boxes, start, count, q, stride, c,
                           cfl, maxHost, maxDev, blockMin);
}

// One SSPRK2 stage of every box (blocks = start[count] of the stage table
// built for chunk depth cz, one of kChunkDepths).
inline void launchBatchStage(int ns, int nv, int scheme, const BoxDev* boxes, const long long* start, int count,
                             long blocks, const double* q, size_t stride, const double* p, const double* c,
                             const double* r, double dtHost, const double* dtDev, int mode, double* out,
                             unsigned long long* failures, hipStream_t stream, int cz = kChunkZ) {
    const unsigned grid = xcdGrid(blocks);
#define RAMD_BATCH_Z(NS_, NV_, H_, Z_)                                                                             \
    hipLaunchKernelGGL((batchTiledStageKernel<NS_, NV_, H_, Z_>), dim3(grid), dim3(kTileX * kTileY), 0, stream,    \
                       boxes, start, count, q, stride, p, c, r, dtHost, dtDev, mode, out, failures)
#define RAMD_BATCH(NS_, NV_, H_)                          \
    do {                                                  \
        if (cz == 2) RAMD_BATCH_Z(NS_, NV_, H_, 2);       \
        else if (cz == 8) RAMD_BATCH_Z(NS_, NV_, H_, 8);  \
        else RAMD_BATCH_Z(NS_, NV_, H_, 4);               \
    } while (0)
#define RAMD_BATCH_SCHEME(NS_, NV_)                  \
    if (scheme == kHllc) RAMD_BATCH(NS_, NV_, true); \
    else RAMD_BATCH(NS_, NV_, false)
    if (ns == 1 && nv == 5) { RAMD_BATCH_SCHEME(1, 5); }
    else if (ns == 1 && nv == 6) { RAMD_BATCH_SCHEME(1, 6); }
    else if (ns == 2 && nv == 6) { RAMD_BATCH_SCHEME(2, 6); }
    else { RAMD_BATCH_SCHEME(2, 7); }
#undef RAMD_BATCH_SCHEME
#undef RAMD_BATCH
#undef RAMD_BATCH_Z
}

}  // namespace ramd
#endif

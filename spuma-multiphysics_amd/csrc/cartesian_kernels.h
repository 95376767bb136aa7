// SPDX-License-Identifier: GPL-3.0-or-later
// Fused structured-box stage kernel (see cartesian.hip for the parity argument).
//
// One thread per cell: the cell's six faces are visited in its CSR order,
// each face's HLL coefficients rebuilt from the two adjacent states, and the
// SSPRK2 update written once.  Workgroups are remapped so that each of the
// eight XCDs sweeps one contiguous z-range: the +/-nx*ny neighbours a cell
// reads were then loaded by the same XCD moments earlier and hit its L2.
#ifndef REACTIVE_AMD_CARTESIAN_KERNELS_H
#define REACTIVE_AMD_CARTESIAN_KERNELS_H
#include <hip/hip_runtime.h>
#include <cstdint>
#include "transport_internal.h"
#include "flux_device.h"

namespace ramd {

struct CartesianArgs {
    int nx, ny, nz;
    double area, invV, waveFactor;
    uint32_t slotOrder[64];
    double normal[6][2][3];
};

inline CartesianArgs makeCartesianArgs(const CartesianInfo& c, double waveFactor) {
    CartesianArgs a;
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

// HLLC = true uses the HLLC option (flux_device.h): rArr holds the recovered
// densities, and a face the HLLC flux rejects keeps its HLL coefficients and
// is counted in failures (each face is seen by both its cells).
template <int NS, int NV, bool HLLC>
__global__ __launch_bounds__(256) void cartesianStageKernel(CartesianArgs a, const double* __restrict__ q,
                                                            size_t stride, const double* __restrict__ pArr,
                                                            const double* __restrict__ cArr,
                                                            const double* __restrict__ rArr, double dtHost,
                                                            const double* __restrict__ dtDev, int mode,
                                                            double* __restrict__ out,
                                                            unsigned long long* __restrict__ failures) {
    const long cells = long(a.nx) * a.ny * a.nz;
    const long blocks = (cells + 255) / 256;
    const long b = xcdBlock(blockIdx.x, blocks);
    const long c = b * 256 + threadIdx.x;
    if (b >= blocks || c >= cells) return;
    const double dt = dtDev ? *dtDev : dtHost;
    if (!(dt > 0)) return;  // an invalid device-side CFL step leaves q untouched
    const long nxy = long(a.nx) * a.ny;
    const int i = int(c % a.nx), j = int((c / a.nx) % a.ny), k = int(c / nxy);
    const int type = (i == 0) | (i == a.nx - 1) << 1 | (j == 0) << 2 | (j == a.ny - 1) << 3 | (k == 0) << 4 |
                     (k == a.nz - 1) << 5;
    const uint32_t order = a.slotOrder[type];
    double qc[NV];
    for (int v = 0; v < NV; ++v) qc[v] = q[size_t(v) * stride + c];
    const double pc = pArr[c], cc = cArr[c];
    const double rc = HLLC ? rArr[c] : 0.0;
    uint32_t rejected = 0;
    double rate[NV];
    for (int v = 0; v < NV; ++v) rate[v] = 0;
    for (int pos = 0; pos < 6; ++pos) {
        const int slot = int((order >> (3 * pos)) & 7u);
        const int ax = slot >> 1, plus = slot & 1;
        const int coord = ax == 0 ? i : (ax == 1 ? j : k);
        const int ext = ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz);
        const bool bnd = plus ? coord == ext - 1 : coord == 0;
        const long step = ax == 0 ? 1 : (ax == 1 ? long(a.nx) : nxy);
        const double* n = a.normal[slot][bnd ? 1 : 0];
        double qn[NV], pn = pc, cn = cc, rn = rc;
        if (bnd) {
            for (int v = 0; v < NV; ++v) qn[v] = qc[v];
        } else {
            const long nb = plus ? c + step : c - step;
            for (int v = 0; v < NV; ++v) qn[v] = q[size_t(v) * stride + nb];
            pn = pArr[nb];
            cn = cArr[nb];
            if (HLLC) rn = rArr[nb];
        }
        // owner (L) and right (R) of this face
        const bool selfLeft = bnd || plus;
        double qL[NV], qR[NV];
        for (int v = 0; v < NV; ++v) {
            qL[v] = selfLeft ? qc[v] : qn[v];
            qR[v] = selfLeft ? qn[v] : qc[v];
        }
        const double pL = selfLeft ? pc : pn, pR = selfLeft ? pn : pc;
        const double aL = a.waveFactor * (selfLeft ? cc : cn), aR = a.waveFactor * (selfLeft ? cn : cc);
        double rhoL = 0, rhoR = 0;
        for (int s = 0; s < NS; ++s) { rhoL += qL[s]; rhoR += qR[s]; }
        double unL = 0, unR = 0;
        for (int d = 0; d < 3; ++d) {
            unL += qL[NS + d] / rhoL * n[d];
            unR += qR[NS + d] / rhoR * n[d];
        }
        const double left = unL - aL < unR - aR ? unL - aL : unR - aR;
        const double right = unL + aL > unR + aR ? unL + aL : unR + aR;
        const double wl = 0 < left ? 0 : left, wr = 0 > right ? 0 : right;
        const double invWave = 1 / (wr - wl);
        double4 w;
        w.x = wr * (unL - wl) * invWave;
        w.y = wl * (wr - unR) * invWave;
        w.z = (wr * pL - wl * pR) * invWave;
        w.w = (wr * pL * unL - wl * pR * unR) * invWave;
        if (HLLC) {
            double uL[3], uR[3];
            for (int d = 0; d < 3; ++d) {
                uL[d] = qL[NS + d] / rhoL;
                uR[d] = qR[NS + d] / rhoR;
            }
            if (!hllcCoefficients(selfLeft ? rc : rn, uL, pL, aL, qL[NS + 3], selfLeft ? rn : rc, uR, pR, aR,
                                  qR[NS + 3], n, w))
                ++rejected;
        }
        const double advectL = w.x, advectR = w.y, pM = w.z, pE = w.w;
        const double sign = selfLeft ? -1.0 : 1.0;
        for (int v = 0; v < NV; ++v) {
            double flux = advectL * qL[v] + advectR * qR[v];
            if (v >= NS && v < NS + 3) flux += (pM - 0.0) * n[v - NS];
            if (v == NS + 3) flux += pE;
            if (v < NS) flux += 0.0;
            else if (v < NS + 3) flux -= 0.0;
            else if (v == NS + 3) flux += 0.0;
            rate[v] += sign * (flux * a.area) * a.invV;
        }
    }
    for (int v = 0; v < NV; ++v) {
        const size_t idx = size_t(v) * stride + c;
        if (mode == 0) out[idx] = qc[v] + dt * rate[v];
        else out[idx] = .5 * out[idx] + .5 * (qc[v] + dt * rate[v]);
    }
    if (HLLC && rejected) atomicAdd(failures, (unsigned long long)rejected);
}

// ---- tiled stage: every face computed once -------------------------------
//
// A workgroup owns a kTileX x kTileY tile of (i, j) columns over kChunkZ
// cells in k and marches up in k.  The +x and +y faces are computed by the
// cell on their minus side and shared through LDS, the +z face by the
// column's own thread (it is the next cell's -z face); a tile's -x / -y edge
// cells compute their -x / -y faces themselves, from the neighbour tile's
// cells.  Velocities are computed once per cell.  Each face is evaluated
// with the per-cell kernel's arithmetic and every cell sums its six faces in
// its CSR order, so the result is the same bit for bit.
#ifndef RAMD_TILE_Z
#define RAMD_TILE_Z 4
#endif
#ifndef RAMD_STAGE_TILED
#define RAMD_STAGE_TILED 1
#endif
constexpr int kTileX = 32, kTileY = 8, kChunkZ = RAMD_TILE_Z;

template <int NS, int NV, bool HLLC>
struct TileCell {
    double q[NV], p, c, rho, r, u[3];  // r: recovered density (HLLC only)
};

template <int NS, int NV, bool HLLC>
__device__ inline void loadTileCell(TileCell<NS, NV, HLLC>& t, const double* __restrict__ q, size_t stride,
                                    const double* __restrict__ pArr, const double* __restrict__ cArr,
                                    const double* __restrict__ rArr, long c) {
    for (int v = 0; v < NV; ++v) t.q[v] = q[size_t(v) * stride + c];
    t.p = pArr[c];
    t.c = cArr[c];
    t.r = HLLC ? rArr[c] : 0.0;
    t.rho = 0;
    for (int s = 0; s < NS; ++s) t.rho += t.q[s];
    for (int d = 0; d < 3; ++d) t.u[d] = t.q[NS + d] / t.rho;
}

// flux * area of one face, L = owner side, R = right side (the per-cell
// kernel's expressions, velocities precomputed)
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
__global__ __launch_bounds__(kTileX * kTileY) void cartesianTiledStageKernel(
    CartesianArgs a, const double* __restrict__ q, size_t stride, const double* __restrict__ pArr,
    const double* __restrict__ cArr, const double* __restrict__ rArr, double dtHost, const double* __restrict__ dtDev,
    int mode, double* __restrict__ out, unsigned long long* __restrict__ failures) {
    using Cell = TileCell<NS, NV, HLLC>;
    __shared__ Cell shared[kTileY][kTileX];
    __shared__ double faceX[kTileY][kTileX][NV], faceY[kTileY][kTileX][NV];
    const int tilesX = (a.nx + kTileX - 1) / kTileX, tilesY = (a.ny + kTileY - 1) / kTileY;
    const int chunks = (a.nz + kChunkZ - 1) / kChunkZ;
    const long blocks = long(tilesX) * tilesY * chunks;
    const long b = xcdBlock(blockIdx.x, blocks);
    if (b >= blocks) return;  // uniform per workgroup
    const double dt = dtDev ? *dtDev : dtHost;
    if (!(dt > 0)) return;  // uniform: an invalid device-side CFL step leaves q untouched
    const int bx = int(b % tilesX), by = int((b / tilesX) % tilesY), bz = int(b / (long(tilesX) * tilesY));
    const int tx = threadIdx.x % kTileX, ty = threadIdx.x / kTileX;
    const int i = bx * kTileX + tx, j = by * kTileY + ty;
    const bool active = i < a.nx && j < a.ny;
    const int k0 = bz * kChunkZ, k1 = min(a.nz, k0 + kChunkZ);
    const long nxy = long(a.nx) * a.ny;
    uint32_t rejected = 0;
    Cell cur, nxt;
    double fzMinus[NV];
    if (active) {
        const long c0 = long(i) + long(a.nx) * j + nxy * k0;
        loadTileCell(cur, q, stride, pArr, cArr, rArr, c0);
        if (k0 == 0) {
            tileFace(a, cur, cur, a.normal[4][1], fzMinus, rejected);
        } else {
            Cell below;
            loadTileCell(below, q, stride, pArr, cArr, rArr, c0 - nxy);
            tileFace(a, below, cur, a.normal[5][0], fzMinus, rejected);
        }
    }
    for (int k = k0; k < k1; ++k) {
        const long c = long(i) + long(a.nx) * j + nxy * k;
        if (active) shared[ty][tx] = cur;
        __syncthreads();
        double fxMinus[NV], fyMinus[NV], fzPlus[NV];
        if (active) {
            double F[NV];
            // +x face (this cell is its owner)
            if (i == a.nx - 1) {
                tileFace(a, cur, cur, a.normal[1][1], F, rejected);
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
                tileFace(a, cur, cur, a.normal[3][1], F, rejected);
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
                tileFace(a, cur, cur, a.normal[5][1], fzPlus, rejected);
            } else {
                loadTileCell(nxt, q, stride, pArr, cArr, rArr, c + nxy);
                tileFace(a, cur, nxt, a.normal[5][0], fzPlus, rejected);
            }
            // -x / -y faces: domain boundary, or a tile edge (neighbour in another tile)
            if (i == 0) {
                tileFace(a, cur, cur, a.normal[0][1], fxMinus, rejected);
            } else if (tx == 0) {
                Cell nb;
                loadTileCell(nb, q, stride, pArr, cArr, rArr, c - 1);
                tileFace(a, nb, cur, a.normal[1][0], fxMinus, rejected);
            }
            if (j == 0) {
                tileFace(a, cur, cur, a.normal[2][1], fyMinus, rejected);
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
                else out[idx] = .5 * out[idx] + .5 * (cur.q[v] + dt * rate[v]);
            }
            for (int v = 0; v < NV; ++v) fzMinus[v] = fzPlus[v];
            cur = nxt;
        }
        __syncthreads();
    }
    if (HLLC && rejected) atomicAdd(failures, (unsigned long long)rejected);
}

inline unsigned tiledGrid(const CartesianArgs& a) {
    const long tiles = long((a.nx + kTileX - 1) / kTileX) * ((a.ny + kTileY - 1) / kTileY) *
                       ((a.nz + kChunkZ - 1) / kChunkZ);
    return xcdGrid(tiles);
}

// FaceSpeeds + Step + the block level of Minimum (reactiveTransportKernels.h:
// 315-342, 529-550; reactiveTransport.cpp:222-232) on the structured box.
// Face speeds are rebuilt per cell and summed in CSR order; invalid cell
// steps map to -1 (absorbing).  maximum comes from maxDev when non-null.
template <int NS>
__global__ __launch_bounds__(256) void cartesianCflKernel(CartesianArgs a, const double* __restrict__ q,
                                                          size_t stride, const double* __restrict__ pArr,
                                                          const double* __restrict__ cArr, double cfl,
                                                          double maxHost, const double* __restrict__ maxDev,
                                                          double* __restrict__ blockMin) {
    (void)pArr;
    const long cells = long(a.nx) * a.ny * a.nz;
    const long blocks = (cells + 255) / 256;
    const long b = xcdBlock(blockIdx.x, blocks);
    if (b >= blocks) return;
    const long c = b * 256 + threadIdx.x;
    double v = __builtin_huge_val();
    if (c < cells) {
        const double maximum = maxDev ? *maxDev : maxHost;
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
            if (bnd) {
                for (int s = 0; s < NS + 3; ++s) qn[s] = qc[s];
            } else {
                const long nb = plus ? c + stepIdx : c - stepIdx;
                for (int s = 0; s < NS + 3; ++s) qn[s] = q[size_t(s) * stride + nb];
                cn = cArr[nb];
            }
            const bool selfLeft = bnd || plus;
            double rhoL = 0, rhoR = 0, unL = 0, unR = 0;
            for (int s = 0; s < NS; ++s) {
                rhoL += selfLeft ? qc[s] : qn[s];
                rhoR += selfLeft ? qn[s] : qc[s];
            }
            for (int d = 0; d < 3; ++d) {
                unL += (selfLeft ? qc[NS + d] : qn[NS + d]) / rhoL * n[d];
                unR += (selfLeft ? qn[NS + d] : qc[NS + d]) / rhoR * n[d];
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
        blockMin[b] = m;
    }
}

inline void launchCartesianCfl(const CartesianArgs& a, int ns, const double* q, size_t stride, const double* p,
                               const double* c, double cfl, double maxHost, const double* maxDev, double* blockMin,
                               hipStream_t stream) {
    const long cells = long(a.nx) * a.ny * a.nz;
    const unsigned grid = xcdGrid((cells + 255) / 256);
    if (ns == 1)
        hipLaunchKernelGGL((cartesianCflKernel<1>), dim3(grid), dim3(256), 0, stream, a, q, stride, p, c, cfl, maxHost,
                           maxDev, blockMin);
    else
        hipLaunchKernelGGL((cartesianCflKernel<2>), dim3(grid), dim3(256), 0, stream, a, q, stride, p, c, cfl, maxHost,
                           maxDev, blockMin);
}

inline bool cartesianSupported(int ns, int nv) {
    return (ns == 1 && (nv == 5 || nv == 6)) || (ns == 2 && (nv == 6 || nv == 7));
}

inline void launchCartesianStage(const CartesianArgs& a, int ns, int nv, int scheme, const double* q, size_t stride,
                                 const double* p, const double* c, const double* r, double dt, const double* dtDev,
                                 int mode, double* out, unsigned long long* failures, hipStream_t stream) {
#if RAMD_STAGE_TILED
    const unsigned grid = tiledGrid(a);
#define RAMD_CART(NS_, NV_, H_)                                                                                    \
    hipLaunchKernelGGL((cartesianTiledStageKernel<NS_, NV_, H_>), dim3(grid), dim3(kTileX * kTileY), 0, stream, a, \
                       q, stride, p, c, r, dt, dtDev, mode, out, failures)
#else
    const long cells = long(a.nx) * a.ny * a.nz;
    const unsigned grid = xcdGrid((cells + 255) / 256);
#define RAMD_CART(NS_, NV_, H_)                                                                                   \
    hipLaunchKernelGGL((cartesianStageKernel<NS_, NV_, H_>), dim3(grid), dim3(256), 0, stream, a, q, stride, p, c, \
                       r, dt, dtDev, mode, out, failures)
#endif
#define RAMD_CART_SCHEME(NS_, NV_)                  \
    if (scheme == kHllc) RAMD_CART(NS_, NV_, true); \
    else RAMD_CART(NS_, NV_, false)
    if (ns == 1 && nv == 5) { RAMD_CART_SCHEME(1, 5); }
    else if (ns == 1 && nv == 6) { RAMD_CART_SCHEME(1, 6); }
    else if (ns == 2 && nv == 6) { RAMD_CART_SCHEME(2, 6); }
    else { RAMD_CART_SCHEME(2, 7); }
#undef RAMD_CART_SCHEME
#undef RAMD_CART
}

}  // namespace ramd
#endif

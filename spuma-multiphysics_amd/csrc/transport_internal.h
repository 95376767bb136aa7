// SPDX-License-Identifier: GPL-3.0-or-later
// Transport handle shared by transport.hip (generic face graphs, ABI) and
// cartesian.hip (structured-box fast path) and step.hip (resident SSPRK2 step).
#ifndef REACTIVE_AMD_TRANSPORT_INTERNAL_H
#define REACTIVE_AMD_TRANSPORT_INTERNAL_H
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>
#include "reactive_transport.h"

namespace ramd {

// Compact face record (64 B): right = neighbour, the owner itself for
// extrapolated boundaries, or cells + fixed for fixed-state boundaries;
// weight = ownerWeight, the owner's share of face interpolation.
struct FaceDev {
    int32_t owner, right, kind, pad;
    double n[3], area, distance, weight;
};
static_assert(sizeof(FaceDev) == 64, "face record");

// Device-side description of one structured box (see CartesianInfo).
struct CartesianArgs {
    int nx, ny, nz;
    double area, invV, waveFactor;
    uint32_t slotOrder[64];
    double normal[6][2][3];
    const int32_t* boundaryKind;  // CartesianInfo::boundaryKind (null: all extrapolated)
    long long fixedBase;
};
// A box of a launch table: its cells are [offset, offset + cells) of the arrays.
struct BoxDev {
    CartesianArgs a;
    unsigned long long offset, cells;
    double volume;
};

// Structured N^3-like box detected from a face list: cell c = i + nx (j + ny k),
// axis-aligned unit normals, one face area, one cell volume, inviscid
// internal (kind 0) faces and boundary faces of any kind: extrapolated (2),
// slip wall (1) or fixed state (3).  The order in which a cell's six faces
// appear in its CSR row (face id order) is a function of the cell's boundary
// type only; slotOrder[type] packs that order (3 bits per position, slots
// 0..5 = -x +x -y +y -z +z), type = 6 boundary flags.
struct CartesianInfo {
    bool valid = false;
    int nx = 0, ny = 0, nz = 0;
    double area = 0, inverseVolume = 0, volume = 0;
    uint32_t slotOrder[64] = {};
    double normal[6][2][3] = {};  // exact normal of [slot][boundary face?]
    // kind of every boundary face by side plane (boundaryPlane): -1
    // extrapolated, -2 slip wall, j >= 0 fixed state j; null when all are
    // extrapolated
    int32_t* boundaryKind = nullptr;
    long long fixedBase = 0;  // array index of fixed state 0 (cells)
    // one-box launch table of the transport's own stages: box, and the
    // stage / CFL block starts {0, stageBlocks, 0, cflBlocks}
    BoxDev* box = nullptr;
    long long* starts = nullptr;
    long stageBlocks = 0, cflBlocks = 0;
};

class Transport {
public:
    std::mutex entry;
    ReactiveTransportConfig cfg;
    size_t all = 0, nBoundary = 0, partitions = 1, bridgeCells = 256;
    int nv = 0, ns = 0, device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    // device arrays
    FaceDev* faces = nullptr;
    double* invV = nullptr;
    uint32_t* rowDev = nullptr;
    int32_t* incidenceDev = nullptr;
    uint32_t* boundaryDev = nullptr;
    double *q = nullptr, *initial = nullptr, *rhsBuffer = nullptr;
    ReactiveTransportState* state = nullptr;
    double* prim = nullptr;            // rho, u0, u1, u2 SoA over cells + fixed
    double4* work = nullptr;           // HLL advectL, advectR, pressureMomentum, pressureEnergy per face
    double* faceSpeed = nullptr;
    double* blockMin = nullptr;
    double *boundaryPartial = nullptr, *boundaryRate = nullptr;
    double* staging = nullptr;         // cell-major host-layout bridge
    double *pressure = nullptr, *sound = nullptr;  // SoA p, c_frozen for the Cartesian path
    double* density = nullptr;                     // SoA recovered density (HLLC on the Cartesian path)
    int scheme = 0;                                // face flux: 0 HLL, 1 HLLC
    unsigned long long* fluxFailures = nullptr;    // faces the HLLC flux rejected since the last check
    // viscous stress, conduction and the WALE closure (generic face graphs only)
    double* gradient = nullptr;     // cells x 9: gradient[9 c + 3 i + j] = d u_i / d x_j
    double* nut = nullptr;          // WALE kinematic eddy viscosity per cell
    double* filterWidth = nullptr;  // WALE filter width cbrt(cell volume) per cell
    double4* visc = nullptr;        // per face: viscous traction (x, y, z), viscous + conductive energy flux (w)
    unsigned* waleError = nullptr;  // set by a WALE viscosity or face that the closure rejects
    bool wale = false;
    double waleCw = 0;
    bool stepAttached = false;      // a resident step was created on this transport
    // resolved interface with surface tension (reactive_transport_set_capillary_v1)
    bool capillary = false, colorReady = false, geometryShadowValid = false;
    double capSigma = 0, capCfl = 0, capEpsilon = 0;
    int capSpecies = 0;
    double* capColor = nullptr;        // cells + fixed
    double* capGradient = nullptr;     // cells x 3
    double* capNormal = nullptr;       // cells x 3
    double* capArea = nullptr;         // |grad colour| per cell
    double* capCurvature = nullptr;    // per cell
    double* capEnergy = nullptr;       // sigma |grad colour| per cell
    double* capLength = nullptr;       // cbrt(V) per cell (host cbrt)
    unsigned* capError = nullptr;
    std::vector<double> colorShadow, fixedColorShadow, energyShadow, curvatureShadow;
    ReactiveCapillaryProfileV1 capillaryProfile{1, sizeof(ReactiveCapillaryProfileV1), 0, 0, 0, 0, 0, 0, 0};
    ReactiveWaleProfileV1 waleProfile{1, sizeof(ReactiveWaleProfileV1), 0, 0, 0, 0};
    CartesianInfo cartesian;
    // versions, tokens, attempts (reactiveTransport.cpp:384-393)
    std::string physicalHash;
    uint64_t attemptId = 0, lastAttempt = 0, tokenStage = 0, residentVersion = 0, highestVersion = 0;
    bool attemptOpen = false, haveResident = false, haveInitial = false;
    double stageDt = 0;
    std::vector<ReactiveTransportState> stateShadow;
    bool stateShadowValid = false;
    // counters
    ReactiveTransportStats stats{};
    ReactiveTransportProfile profile{};
    ReactiveTransportDeviceProfile deviceProfile{};
    ReactiveTransportProfileV21 cost{};
    uint64_t peakBytes = 0, deviceCopies = 0, synchronizations = 0, memcpyCalls = 0;
    char error[1024] = {};

    Transport(int backend, const ReactiveTransportConfig& c, const double* volumes, const ReactiveTransportFace* faces,
              const double* fixedQ, const ReactiveTransportState* fixedStates,
              const ReactiveTransportOptionsV2* options = nullptr,
              const ReactiveTransportOptionsV21* options21 = nullptr);
    ~Transport();
    void* allocateBytes(size_t bytes);
    template <class T> T* allocate(size_t n) { return static_cast<T*>(allocateBytes(n * sizeof(T))); }
    void uploadBytes(void* dst, const void* src, size_t bytes);
    void downloadBytes(void* dst, const void* src, size_t bytes);
    template <class T> void upload(T* dst, const T* src, size_t n) { uploadBytes(dst, src, n * sizeof(T)); }
    void finish();
    void launched();
    static void validateStates(const ReactiveTransportState* s, size_t count);
    void launchPack(const double* aos, size_t cells, double* soa, size_t offset);
    void launchUnpack(const double* soa, double* hostOut);
    void uploadQ(const double* hostQ);
    void uploadState(const ReactiveTransportState* states);
    void primitives();
    // true where the face flux carries viscous, conductive or WALE terms
    bool viscous() const { return cfg.viscosity > 0 || cfg.conductivity > 0 || wale; }
    // true where the face records carry a traction / energy part (visc)
    bool faceRecords() const { return viscous() || capillary; }
    void installWale(const ReactiveWaleOptionsV1& options);
    void installCapillary(const ReactiveCapillaryOptionsV1& options);
    void capillaryGeometry(const double* cellColor, const double* fixedColor, double* energy, double* curvature,
                           double* normal);
    void launchGradients(bool withWale);
    void buildWale();
    void launchFaces();
    void prepareResident(const ReactiveTransportState* states, bool transport);
    void launchFaceSpeeds();
    void gather();
    void stableStep(double cfl, double maximum, double* dt);
    void advance(double dt, int stage, double* hostQ, double* boundary);
    void rhs(double* hostRhs, double* boundary);
    void uploadPrimitives(const ReactiveTransportPrimitive* primitive, const ReactiveTransportState* states);
    // cartesian.hip
    CartesianInfo detectCartesian(const ReactiveTransportFace* hostFaces, const double* volumes,
                                  const std::vector<uint32_t>& row, const std::vector<int32_t>& incidence);
    void cartesianAdvance(double dt, int stage);
    // device-resident step support (step.hip): boundary flux rates of q
    // (variable stride `stride`, `all` unless the arrays hold a batch of
    // boxes) with the face coefficients rebuilt from q, p, c (and the
    // recovered density r for HLLC) on the fly; HLLC rejections count into failures
    void boundaryRatesFrom(const double* qv, const double* p, const double* c, const double* r, double* out,
                           unsigned long long* failures, hipStream_t s, size_t stride);
    // throws (and clears the counter) if the HLLC flux rejected a face
    void checkFluxFailures();
};

}  // namespace ramd
#endif

"""ctypes mirrors of the C ABI structures (include/reactive_transport.h,
include/reactive_gpu_hem.h) plus numpy dtypes of the same layouts."""
import ctypes as C
import numpy as np


class ThermoState(C.Structure):
    """ReactiveThermoState (reactiveThermo.h:11-19, 176 B)."""
    _fields_ = [("p", C.c_double), ("T", C.c_double), ("rho", C.c_double), ("e", C.c_double),
                ("entropy", C.c_double), ("soundFrozen", C.c_double), ("soundEquilibrium", C.c_double),
                ("volumeResidual", C.c_double), ("energyResidual", C.c_double),
                ("chemicalResidual", C.c_double), ("alphaGas", C.c_double),
                ("alphaLiquid", C.c_double * 2), ("rhoGas", C.c_double), ("rhoLiquid", C.c_double * 2),
                ("liquidMass", C.c_double * 2), ("gasMass", C.c_double), ("cp", C.c_double),
                ("cv", C.c_double), ("activeLiquids", C.c_int), ("iterations", C.c_int)]


STATE_DTYPE = np.dtype([("p", "<f8"), ("T", "<f8"), ("rho", "<f8"), ("e", "<f8"), ("entropy", "<f8"),
                        ("soundFrozen", "<f8"), ("soundEquilibrium", "<f8"), ("volumeResidual", "<f8"),
                        ("energyResidual", "<f8"), ("chemicalResidual", "<f8"), ("alphaGas", "<f8"),
                        ("alphaLiquid", "<f8", (2,)), ("rhoGas", "<f8"), ("rhoLiquid", "<f8", (2,)),
                        ("liquidMass", "<f8", (2,)), ("gasMass", "<f8"), ("cp", "<f8"), ("cv", "<f8"),
                        ("activeLiquids", "<i4"), ("iterations", "<i4")])
assert C.sizeof(ThermoState) == 176 and STATE_DTYPE.itemsize == 176


class HemProfile(C.Structure):
    """ReactiveGpuHemProfileV1 (reactiveGpuHem.h:14-21, 144 B)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32)] + [
        (n, C.c_uint64) for n in ("batches", "submitted", "succeeded", "cpuFallbacks", "deviceFailures",
                                  "phaseEvaluations", "residualEvaluations", "flashCandidates",
                                  "stableCandidates", "analyticJacobians", "finiteDifferenceJacobians",
                                  "transferBytes", "hostBytes", "deviceBytes")] + [
        (n, C.c_double) for n in ("wallSeconds", "kernelSeconds", "copySeconds")]

    @classmethod
    def fresh(cls):
        p = cls()
        p.abiVersion = 1
        p.structBytes = C.sizeof(cls)
        return p


assert C.sizeof(HemProfile) == 144


class CapillaryInput(C.Structure):
    _fields_ = [("color", C.c_double), ("pressureJump", C.c_double), ("equilibrium", C.c_int)]


class HemOutput(C.Structure):
    _fields_ = [("state", ThermoState), ("success", C.c_int), ("reserved", C.c_int)]


assert C.sizeof(CapillaryInput) == 24 and C.sizeof(HemOutput) == 184
HEM_OUTPUT_DTYPE = np.dtype([("state", STATE_DTYPE), ("success", "<i4"), ("reserved", "<i4")])
CAPILLARY_DTYPE = np.dtype([("color", "<f8"), ("pressureJump", "<f8"), ("equilibrium", "<i4"), ("pad", "<i4")])
assert CAPILLARY_DTYPE.itemsize == 24

# flash arithmetic (REACTIVE_AMD_FLASH_FP64 / _MIXED, include/reactive_amd.h)
FP64, MIXED = 0, 1


class Face(C.Structure):
    """ReactiveTransportFace (reactiveTransport.h:10-14, 80 B)."""
    _fields_ = [("owner", C.c_int64), ("neighbour", C.c_int64), ("fixed", C.c_int64),
                ("normal", C.c_double * 3), ("area", C.c_double), ("distance", C.c_double),
                ("ownerWeight", C.c_double), ("kind", C.c_int)]


FACE_DTYPE = np.dtype([("owner", "<i8"), ("neighbour", "<i8"), ("fixed", "<i8"), ("normal", "<f8", (3,)),
                       ("area", "<f8"), ("distance", "<f8"), ("ownerWeight", "<f8"), ("kind", "<i4"),
                       ("pad", "<i4")])
assert C.sizeof(Face) == 80 and FACE_DTYPE.itemsize == 80


class TransportState(C.Structure):
    """ReactiveTransportState (reactiveTransport.h:15-17, 56 B)."""
    _fields_ = [(n, C.c_double) for n in ("p", "T", "rho", "cv", "sound", "gasMass", "dilatation")]


TSTATE_DTYPE = np.dtype([(n, "<f8") for n in ("p", "T", "rho", "cv", "sound", "gasMass", "dilatation")])


class Config(C.Structure):
    """ReactiveTransportConfig (reactiveTransport.h:18-23, 88 B)."""
    _fields_ = [("cells", C.c_size_t), ("species", C.c_size_t), ("variables", C.c_size_t),
                ("faces", C.c_size_t), ("fixed", C.c_size_t), ("viscosity", C.c_double),
                ("conductivity", C.c_double), ("diffusivity", C.c_double), ("waveFactor", C.c_double),
                ("maxBytes", C.c_double), ("mechanical", C.c_int)]


class Token(C.Structure):
    _fields_ = [("attemptId", C.c_uint64), ("stageId", C.c_uint64), ("contentVersion", C.c_uint64)]


class OptionsV2(C.Structure):
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("bridgeCells", C.c_size_t),
                ("recomputeGas", C.c_int), ("physicalModelHash", C.c_char * 65)]


class OptionsV21(C.Structure):
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("slotBytes", C.c_size_t),
                ("pinnedBudgetBytes", C.c_size_t), ("bridgeCellsOverride", C.c_size_t),
                ("slots", C.c_uint32), ("blockThreads", C.c_uint32), ("recomputeGas", C.c_int),
                ("detailedGasCounters", C.c_int), ("physicalModelHash", C.c_char * 65)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("allocatedBytes", "uploadedBytes", "downloadedBytes",
                                          "kernelLaunches", "stages", "stepQueries")]


class Primitive(C.Structure):
    _fields_ = [("rho", C.c_double), ("u", C.c_double * 3)]


PRIMITIVE_DTYPE = np.dtype([("rho", "<f8"), ("u", "<f8", (3,))])


class WaleOptions(C.Structure):
    """ReactiveWaleOptionsV1 (reactiveTurbulence.h:12-15)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("Cw", C.c_double)]

    @classmethod
    def make(cls, cw):
        return cls(abiVersion=1, structBytes=C.sizeof(cls), Cw=cw)


class CapillaryOptions(C.Structure):
    """ReactiveCapillaryOptionsV1 (reactiveTransport.h:38-42)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("sigma", C.c_double),
                ("capillaryCfl", C.c_double), ("geometryEpsilon", C.c_double), ("condensableSpecies", C.c_int64)]

    @classmethod
    def make(cls, sigma, capillary_cfl=0.5, epsilon=1e-6, species=0):
        return cls(abiVersion=1, structBytes=C.sizeof(cls), sigma=sigma, capillaryCfl=capillary_cfl,
                   geometryEpsilon=epsilon, condensableSpecies=species)


class CapillaryProfile(C.Structure):
    """ReactiveCapillaryProfileV1 (reactiveTransport.h:43-47)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("geometryBuilds", C.c_uint64),
                ("geometryCells", C.c_uint64), ("geometryKernels", C.c_uint64), ("faceFluxBuilds", C.c_uint64),
                ("colorUploadBytes", C.c_uint64), ("geometryDownloadBytes", C.c_uint64),
                ("capillaryWorkspaceBytes", C.c_uint64)]

    @classmethod
    def fresh(cls):
        return cls(abiVersion=1, structBytes=C.sizeof(cls))


class WaleProfile(C.Structure):
    """ReactiveWaleProfileV1 (reactiveTurbulence.h:16-19)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("gradientBuilds", C.c_uint64),
                ("viscosityBuilds", C.c_uint64), ("cellsEvaluated", C.c_uint64), ("workspaceBytes", C.c_uint64)]

    @classmethod
    def fresh(cls):
        return cls(abiVersion=1, structBytes=C.sizeof(cls))


class Profile(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("conservedUploads", "conservedUploadBytes", "conservedDownloadBytes",
                                          "stateUploadBytes", "primitiveUploadBytes", "gasUploadBytes",
                                          "boundaryPartitions", "residentStages")]


class DeviceProfile(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("gasPropertyBuilds", "gasPropertyCells", "partitionUploadBytes",
                                          "thermoTableBytes", "gasStatusChecks", "cflFaceLaunches",
                                          "transportFaceLaunches", "faceWorkspaceBytes")]


class MemoryV2(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("liveBytes", "peakBytes", "bridgeBytes", "conservedWorkspaceBytes",
                                          "gasWorkspaceBytes", "rhsWorkspaceBytes", "deviceToDeviceBytes",
                                          "synchronizations")]


class ProfileV21(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("layoutTiles", "layoutLaunches", "memcpyCalls", "payloadBytes",
                                          "nasaValidationEvaluations", "nasaFaceEvaluations",
                                          "temperatureBasisBuilds", "pinnedBytes", "deviceStagingBytes",
                                          "temperatureBasisBytes", "eventWaits", "slotCells", "slots",
                                          "blockThreads", "failedCalls")]


assert C.sizeof(Config) == 88 and C.sizeof(OptionsV2) == 88 and C.sizeof(OptionsV21) == 120
assert C.sizeof(Token) == 24 and C.sizeof(TransportState) == 56 and C.sizeof(Primitive) == 32
assert PRIMITIVE_DTYPE.itemsize == 32 and C.sizeof(WaleOptions) == 16 and C.sizeof(WaleProfile) == 40


class StepReport(C.Structure):
    """ReactiveAmdStepReport (include/reactive_amd.h)."""
    _fields_ = [("abiVersion", C.c_uint32), ("structBytes", C.c_uint32), ("steps", C.c_uint64),
                ("flashFailures", C.c_uint64 * 2), ("cflViolations", C.c_uint64), ("phaseEvaluations", C.c_uint64),
                ("dt", C.c_double), ("recheckDt", C.c_double), ("totals", C.c_double * 32),
                ("boundaryIntegral", C.c_double * 32), ("seconds", C.c_double), ("stageSeconds", C.c_double * 2),
                ("flashSeconds", C.c_double * 2), ("cflSeconds", C.c_double), ("fluxFailures", C.c_uint64),
                ("singlePhaseEvaluations", C.c_uint64), ("searchSeconds", C.c_double),
                ("queuedCells", C.c_uint64), ("tileDepth", C.c_uint32), ("flashBlock", C.c_uint32),
                ("stateBins", C.c_uint32), ("reserved", C.c_uint32), ("retries", C.c_uint64),
                ("reusedCells", C.c_uint64)]

    @classmethod
    def fresh(cls):
        r = cls()
        r.abiVersion = 1
        r.structBytes = C.sizeof(cls)
        return r

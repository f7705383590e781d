"""ctypes mirror of include/mara_gpu.h: status codes, enumerations and the
two plain structs of the C ABI.  No library is loaded here."""
import ctypes

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_GUARD_TOO_SMALL = 2
ERR_UNSUPPORTED = 3
ERR_HIP = 4
ERR_INVERSION_FAILURE = 5
ERR_SUPERLUMINAL = 6
ERR_BAD_EIGENVALUES = 7
ERR_NO_DEVICE = 8
ERR_CAPACITY = 9

LAW_HYDRO, LAW_MHD, LAW_RMHD = 0, 1, 2
RECON_PCM, RECON_PLM, RECON_WENO5 = 0, 1, 2
RIEMANN_HLLE, RIEMANN_HLLC = 0, 1
BC_PERIODIC, BC_OUTFLOW, BC_REFLECTING = 0, 1, 2
PREC_FP64, PREC_MIXED, PREC_FP64_FAST = 0, 1, 2

LAWS = {"hydro": LAW_HYDRO, "mhd": LAW_MHD, "rmhd": LAW_RMHD}
RECONS = {"pcm": RECON_PCM, "plm": RECON_PLM, "weno5": RECON_WENO5}
RIEMANNS = {"hlle": RIEMANN_HLLE, "hllc": RIEMANN_HLLC}
BCS = {"periodic": BC_PERIODIC, "outflow": BC_OUTFLOW, "reflecting": BC_REFLECTING}
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED, "fp64_fast": PREC_FP64_FAST}
GEOM_CARTESIAN = 0
GEOM_SPHERICAL = 1
GEOMETRIES = {"cartesian": GEOM_CARTESIAN, "spherical": GEOM_SPHERICAL}


def num_fields(law):
    """getNumConserved: 5 for NewtonianHydro, 8 for the MHD laws."""
    return 5 if law == LAW_HYDRO else 8


def stencil_size(recon):
    """IntercellFluxScheme::getStencilSize: PCM 1, PLM 2, WENO5 3."""
    return {RECON_PCM: 1, RECON_PLM: 2, RECON_WENO5: 3}[recon]


class MaraDesc(ctypes.Structure):
    _fields_ = [
        ("law", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("pressure_floor", ctypes.c_double),
        ("recon", ctypes.c_int),
        ("plm_theta", ctypes.c_double),
        ("riemann", ctypes.c_int),
        ("bc", ctypes.c_int),
        ("cells", ctypes.c_int * 3),
        ("guards", ctypes.c_int * 3),
        ("lower", ctypes.c_double * 3),
        ("upper", ctypes.c_double * 3),
        ("rk_order", ctypes.c_int),
        ("field_ct", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("cooling_rate", ctypes.c_double),
        ("geometry", ctypes.c_int),
        ("log_radial", ctypes.c_int),
    ]


class MaraShape(ctypes.Structure):
    """mara_shape: launch shape of the stage kernels (pencil length, z planes
    per fused work-group, flux-array work-group rows or 2-D MHD tile height)."""
    _fields_ = [
        ("pencil", ctypes.c_int),
        ("zplanes", ctypes.c_int),
        ("rows", ctypes.c_int),
    ]

    def as_dict(self):
        return {"pencil": self.pencil, "zplanes": self.zplanes, "rows": self.rows}


PENCIL_SHAPES = (128, 256, 512)
ZPLANE_SHAPES = (4, 8, 16, 32)  # the autotuner's fixed candidates; set_shape takes 1 .. ZPLANES_MAX
ZPLANES_MAX = 1024
ROW_SHAPES = (1, 2, 4)
MHD_ROWS_DEFAULT = 2  # fused 2-D MHD tiles of 64 x 8 zones


def make_shape(pencil=256, zplanes=16, rows=4):
    return MaraShape(pencil, zplanes, rows)


class MaraReport(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("c2p_failures", ctypes.c_longlong),
        ("first_failure", ctypes.c_int * 3),
        ("floor_applied", ctypes.c_longlong),
        ("eigen_failures", ctypes.c_longlong),
        ("superluminal", ctypes.c_longlong),
    ]

    def as_dict(self):
        return {
            "status": self.status,
            "c2p_failures": self.c2p_failures,
            "first_failure": tuple(self.first_failure),
            "floor_applied": self.floor_applied,
            "eigen_failures": self.eigen_failures,
            "superluminal": self.superluminal,
        }


def make_desc(law="hydro", gamma=5.0 / 3.0, pressure_floor=-1.0, recon="plm",
              plm_theta=1.5, riemann="hlle", bc="periodic", cells=(128, 1, 1),
              guards=None, lower=(0.0, 0.0, 0.0), upper=(1.0, 1.0, 1.0),
              rk_order=2, field_ct=1, precision="fp64", cooling_rate=0.0, geometry="cartesian",
              log_radial=False):
    """Build a MaraDesc from names.  guards defaults to the scheme's stencil
    size on every axis with more than one cell (the reference programs'
    boundary shapes, src/Problems.cpp:155-156)."""
    d = MaraDesc()
    d.law = LAWS[law] if isinstance(law, str) else law
    d.gamma = gamma
    d.pressure_floor = pressure_floor
    d.recon = RECONS[recon] if isinstance(recon, str) else recon
    d.plm_theta = plm_theta
    d.riemann = RIEMANNS[riemann] if isinstance(riemann, str) else riemann
    d.bc = BCS[bc] if isinstance(bc, str) else bc
    cells = tuple(cells) + (1,) * (3 - len(cells))
    if guards is None:
        g = stencil_size(d.recon)
        guards = tuple(g if n > 1 else 0 for n in cells)
    for a in range(3):
        d.cells[a] = cells[a]
        d.guards[a] = guards[a]
        d.lower[a] = lower[a]
        d.upper[a] = upper[a]
    d.rk_order = rk_order
    d.field_ct = field_ct
    d.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    d.cooling_rate = cooling_rate
    d.geometry = GEOMETRIES[geometry] if isinstance(geometry, str) else geometry
    d.log_radial = 1 if log_radial else 0
    return d


def array_shape(desc):
    """Shape of MeshData::P for a descriptor: (n1+2g1, n2+2g2, n3+2g3, nq)."""
    return tuple(desc.cells[a] + 2 * desc.guards[a] for a in range(3)) + (desc_num_fields(desc),)


def desc_num_fields(desc):
    """Fields of a descriptor's state: num_fields(law), plus the auxiliary
    one of NewtonianMHD with a positive cooling rate."""
    return 9 if desc.law == LAW_MHD and desc.cooling_rate > 0 else num_fields(desc.law)

"""ctypes mirror of include/mahyco.h (config struct, field ids, status codes).

The C header is the contract; this module only restates its data layout for
Python callers (tests, bench, the library loader).  A CPU test checks that
sizeof/offsets agree with the compiled library.
"""
import ctypes

NBMATMAX = 3

MHY_OK = 0
MHY_ERR_ARG = -1
MHY_ERR_IO = -2
MHY_ERR_DEVICE = -3
MHY_ERR_ENERGY = -4
MHY_ERR_NAN = -5
MHY_ERR_UNKNOWN_CASE = -6
MHY_ERR_FINISHED = -7
MHY_ERR_DT_MIN = -8

LIM = {"minmod": 300, "superBee": 301, "vanLeer": 302, "minmodG": 1300, "superBeeG": 1301,
       "vanLeerG": 1302, "arithmeticG": 1303, "ultrabeeG": 1304}
EOS_VOID, EOS_PERFECT_GAS = 100, 101
SCHEMA_EUCCLHYD, SCHEMA_VNR = 2000, 2001

CASE = {"UnitTestCase": 0, "SedovTestCase": 1, "TriplePoint": 2, "SodCaseX": 4, "SodCaseY": 5,
        "NohTestCase": 6, "AdvectionX": 7, "AdvectionY": 8, "AdvectionVitX": 9, "AdvectionVitY": 10,
        "BiUnitTestCase": 11, "BiSedovTestCase": 12, "BiTriplePoint": 13, "BiShockBubble": 14,
        "BiSodCaseX": 15, "BiSodCaseY": 16, "BiNohTestCase": 17, "BiAdvectionX": 18,
        "BiAdvectionY": 19, "BiAdvectionVitX": 20, "BiAdvectionVitY": 21,
        "RiemannPerturbed": 100, "PointExplosion": 101, "TripleLayout": 102, "Riemann": 103}

# field ids
CELL_DENSITY, CELL_PRESSURE, CELL_ENERGY, CELL_PSEUDO, CELL_SOUND_SPEED = 0, 1, 2, 3, 4
CELL_DIVU, CELL_MASS, CELL_MIXTE, CELL_PURE = 5, 6, 7, 8
MAT_FRACVOL, MAT_MASS_FRACTION, MAT_DENSITY, MAT_ENERGY = 16, 17, 18, 19
MAT_PRESSURE, MAT_PSEUDO, MAT_MASS, MAT_SOUND_SPEED = 20, 21, 22, 23
MAT_STRIDE = 16
NODE_VELOCITY_X, NODE_VELOCITY_Y, NODE_MASS = 64, 65, 66
NODE_COORD_X, NODE_COORD_Y, NODE_OUT_VX, NODE_OUT_VY = 67, 68, 69, 70

CELL_FIELDS = {"density": CELL_DENSITY, "pressure": CELL_PRESSURE, "energy": CELL_ENERGY,
               "pseudo": CELL_PSEUDO, "sound_speed": CELL_SOUND_SPEED, "divu": CELL_DIVU,
               "cell_mass": CELL_MASS, "mixte": CELL_MIXTE, "pure": CELL_PURE,
               # Eucclhyd only
               "cell_velocity_x": 9, "cell_velocity_y": 10, "cell_out_vx": 11, "cell_out_vy": 12}
MAT_FIELDS = {"fracvol": MAT_FRACVOL, "mass_fraction": MAT_MASS_FRACTION,
              "density_env": MAT_DENSITY, "energy_env": MAT_ENERGY, "pressure_env": MAT_PRESSURE,
              "pseudo_env": MAT_PSEUDO, "mass_env": MAT_MASS, "sound_speed_env": MAT_SOUND_SPEED}
NODE_FIELDS = {"velocity_x": NODE_VELOCITY_X, "velocity_y": NODE_VELOCITY_Y, "node_mass": NODE_MASS,
               "coord_x": NODE_COORD_X, "coord_y": NODE_COORD_Y, "out_vx": NODE_OUT_VX,
               "out_vy": NODE_OUT_VY}


def mat_field(base, k):
    return base + MAT_STRIDE * k


class MhyConfig(ctypes.Structure):
    _fields_ = [
        ("test_case", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("final_time", ctypes.c_double), ("output_time", ctypes.c_double),
        ("deltat_init", ctypes.c_double), ("deltat_min", ctypes.c_double),
        ("max_time_iterations", ctypes.c_int),
        ("nbmat", ctypes.c_int),
        ("eos", ctypes.c_int * NBMATMAX),
        ("gammap", ctypes.c_double * NBMATMAX),
        ("gamma", ctypes.c_double),
        ("schema", ctypes.c_int),
        ("space_order", ctypes.c_int),
        ("avec_projection", ctypes.c_int),
        ("projection_conservative", ctypes.c_int),
        ("threshold", ctypes.c_double),
        ("projection_order", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("pp", ctypes.c_int),
        ("pp_mixte", ctypes.c_int),
        ("limiter_pure", ctypes.c_int),
        ("pp_complet", ctypes.c_int),
        ("pseudo_centree", ctypes.c_int),
        ("methode_flux_masse", ctypes.c_int),
        ("sans_lagrange", ctypes.c_int),
        ("abort_on_energy", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("seed", ctypes.c_ulonglong),
        ("param", ctypes.c_double * 8),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class MhyTimeState(ctypes.Structure):
    _fields_ = [("t_n", ctypes.c_double), ("t_nplus1", ctypes.c_double),
                ("deltat_n", ctypes.c_double), ("deltat_nplus1", ctypes.c_double),
                ("iteration", ctypes.c_int), ("x_then_y_n", ctypes.c_int),
                ("finished", ctypes.c_int), ("status", ctypes.c_int)]

"""Writes a small Alya solidz deck of a Nitinol beam (files of our own, in
the reference's input format) that uses the Neumann keywords: face codes
with a follower pressure (CODES, BOUNDARIES fixbo 3 + MODULE_FUNCTION,
FOLLOWER_LOADS: ON, TRANSIENT conditions) and nodal forces
(FORCE: FIELD=1 with a FIELDS section)."""
import os

import numpy as np

from alya_sma_amd import problems as P

NX, NY, NZ, LX, LY, LZ = 8, 2, 2, 4.0, 1.0, 1.0
T1 = 4.0e-5
TIP_FORCE = (0.0, 0.0, -3.0)
WITNESS = [(LX, LY, LZ), (0.0, 0.5 * LY, 0.5 * LZ), (1.3, 0.37, 0.61), (LX + 1.0, 0.5, 0.5)]
WITNESS_VARS = ["DISPX", "DISPY", "DISPZ", "FRXIX", "SIGXX", "SIGXY"]
PRESSURE = -5.0


def write_beam_deck(d, name="beam"):
    coord, lnods = P.box_mesh_np(NX, NY, NZ, LX, LY, LZ)
    faces, elem = P.boundary_faces(coord, lnods)
    centre = coord[faces].mean(axis=1)
    code = np.zeros(len(faces), np.int64)
    code[np.abs(centre[:, 0]) < 1e-9] = 1                      # clamped end
    code[np.abs(centre[:, 2] - LZ) < 1e-9] = 2                 # top: pressure
    tip = int(np.argmin(np.sum((coord - [LX, LY, LZ]) ** 2, axis=1)))
    w = lambda n, text: open(os.path.join(d, f"{name}.{n}"), "w").write(text)
    w("dat", f"""RUN_DATA
  ALYA: {name}
END_RUN_DATA
PROBLEM_DATA
  TIME_COUPLING:      GLOBAL, FROM_CRITICAL
  TIME_INTERVAL:      0.0, {T1}
  NUMBER_OF_STEPS=    1000000
END_PROBLEM_DATA
""")
    w("dom.dat", f"""DIMENSIONS
  NODAL_POINTS      = {len(coord)}
  ELEMENTS          = {len(lnods)}
  SPACE_DIMENSIONS  = 3
  TYPES_OF_ELEMENTS = HEX08
  BOUNDARIES        = {len(faces)}
  MATERIALS         = 1
  FIELDS            = 1
    FIELD = 1, DIMENSION = 3, NODES
  END_FIELDS
END_DIMENSIONS
GEOMETRY
  INCLUDE {name}.geo.dat
END_GEOMETRY
SETS
  INCLUDE {name}.set.dat
END_SETS
BOUNDARY_CONDITIONS
  INCLUDE {name}.fix.dat
END_BOUNDARY_CONDITIONS
FIELDS
  FIELD = 1
   {tip + 1} {TIP_FORCE[0]} {TIP_FORCE[1]} {TIP_FORCE[2]}
  END_FIELD
END_FIELDS
""")
    geo = ["ELEMENTS"] + [f" {e + 1} " + " ".join(str(n + 1) for n in row) for e, row in enumerate(lnods)]
    geo += ["END_ELEMENTS", "COORDINATES"]
    geo += [f" {p + 1} {x:.17g} {y:.17g} {z:.17g}" for p, (x, y, z) in enumerate(coord)]
    geo += ["END_COORDINATES", "BOUNDARIES, ELEMENTS"]
    geo += [f" {b + 1} " + " ".join(str(n + 1) for n in row) + f" {elem[b] + 1}"
            for b, row in enumerate(faces)]
    geo += ["END_BOUNDARIES"]
    w("geo.dat", "\n".join(geo) + "\n")
    w("fix.dat", "ON_BOUNDARIES\n" + "".join(f" {b + 1} {c}\n" for b, c in enumerate(code) if c)
      + "END_ON_BOUNDARIES\n")
    w("set.dat", f"NODE\n {tip + 1} 1\nEND_NODE\n")
    # witness points (given in cm with SCALE 10 -> mm): the tip node, the
    # centre of the clamped face, an interior point, one outside the beam
    w("ker.dat", "WITNESS_POINTS, NUMBER=4, SCALE : XSCALE= 10.0 YSCALE= 10.0 ZSCALE= 10.0\n"
      + "".join(f"  {x / 10:.17g} {y / 10:.17g} {z / 10:.17g}\n" for x, y, z in WITNESS)
      + "END_WITNESS_POINTS\n")
    props = " ".join(f"{x:g}" for x in P.NITINOL[:17])
    w("sld.dat", f"""PHYSICAL_PROBLEM
  PROBLEM_DEFINITION
    TEMPORAL_DERIVATIVES: DYNAMIC
  END_PROBLEM_DEFINITION
  PROPERTIES
    MATERIAL = 1
    DENSITY = {P.RHO_NITINOL}
    CONSTITUTIVE_MODEL: NITINOL \\
       {props}
    FORCE: FIELD=1
  END_PROPERTIES
END_PHYSICAL_PROBLEM
OUTPUT_&_POST_PROCESS
  WITNESS_POINTS
    DISPX
    DISPY
    DISPZ
    FRXIX
    SIGXX
    SIGXY
  END_WITNESS_POINTS
END_OUTPUT_&_POST_PROCESS
NUMERICAL_TREATMENT
  TIME_TREATMENT:      EXPLICIT
  TIME_INTEGRATION:    TCHAMWA_WIELGOSZ, PHITW=1.033
  SAFETY_FACTOR=       0.5
END_NUMERICAL_TREATMENT
BOUNDARY_CONDITIONS, TRANSIENT
  PARAMETERS
    FOLLOWER_LOADS: ON
  END_PARAMETERS
  CODES, NODES
    1      111  0.0 0.0 0.0
    1 & 2  111  0.0 0.0 0.0
  END_CODES
  CODES, BOUNDARIES
    2  3  0.0 0.0 0.0, MODULE_FUNCTION= 1
  END_CODES
  FUNCTIONS
    TOTAL_NUMBER = 1
    CONDITION
      FUNCTION_NUMBER: 1
      TIME_SHAPE:      DISCRETE, SMOOTH
      SHAPE_DEFINITION
        3
        0.0      0.0 0.0 0.0
        {T1 / 2} {PRESSURE} 0.0 0.0
        {T1}     0.0 0.0 0.0
      END_SHAPE_DEFINITION
    END_CONDITION
  END_FUNCTIONS
END_BOUNDARY_CONDITIONS
""")
    return os.path.join(d, f"{name}.dat"), dict(coord=coord, lnods=lnods, faces=faces, code=code,
                                                tip=tip)

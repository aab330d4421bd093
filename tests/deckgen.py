"""Writes small OGS-5 decks for the deck-reader tests: a vertical (x-z)
section of quads with two material groups, LIQUID_FLOW feeding
HEAT_TRANSPORT, polyline and point conditions and three kinds of output."""
import os

NX, NZ, LX, LZ = 8, 4, 80.0, 20.0


def quad_xz_msh():
    lines = ["#FEM_MSH", " $PCS_TYPE", "  NO_PCS", " $NODES", f"  {(NX + 1) * (NZ + 1)}"]
    for k in range(NZ + 1):
        for i in range(NX + 1):
            lines.append(f"  {k * (NX + 1) + i} {i * LX / NX} 0 {k * LZ / NZ}")
    lines += [" $ELEMENTS", f"  {NX * NZ}"]
    e = 0
    for k in range(NZ):
        for i in range(NX):
            n0 = k * (NX + 1) + i
            group = 0 if i < NX // 2 else 1
            lines.append(f"  {e} {group} quad {n0} {n0 + 1} {n0 + NX + 2} {n0 + NX + 1}")
            e += 1
    lines.append("#STOP")
    return "\n".join(lines) + "\n"


FILES = {
    ".pcs": """#PROCESS
 $PCS_TYPE
  LIQUID_FLOW
#PROCESS
 $PCS_TYPE
  HEAT_TRANSPORT
#STOP
""",
    ".num": """#NUMERICS
 $PCS_TYPE
  LIQUID_FLOW
 $LINEAR_SOLVER
; method error_tolerance max_iterations theta precond storage
  5      1 1.e-13 5000 1.0 1 2
 $ELE_GAUSS_POINTS
  3
#NUMERICS
 $PCS_TYPE
  HEAT_TRANSPORT
 $LINEAR_SOLVER
  2      1 1.e-13 5000 0.5 100 2
#STOP
""",
    ".tim": """#TIME_STEPPING
 $PCS_TYPE
  LIQUID_FLOW
 $TIME_UNIT
  DAY
 $TIME_START
  0
 $TIME_END
  3
 $TIME_STEPS
  3 1.0
#TIME_STEPPING
 $PCS_TYPE
  HEAT_TRANSPORT
 $TIME_UNIT
  DAY
 $TIME_START
  0
 $TIME_END
  3
 $TIME_STEPS
  3 1.0
#STOP
""",
    ".gli": """#POINTS
 0 0 0 0 $NAME PNT0
 1 0 0 20
 2 80 0 0
 3 80 0 20
 4 40 0 10 $NAME WELL
 5 61 0 9 $NAME OBS
#POLYLINE
 $NAME
  LEFT
 $POINTS
  0
  1
#POLYLINE
 $NAME
  RIGHT
 $POINTS
  2
  3
#STOP
""",
    ".bc": """#BOUNDARY_CONDITION
 $PCS_TYPE
  LIQUID_FLOW
 $PRIMARY_VARIABLE
  PRESSURE1
 $GEO_TYPE
  POLYLINE LEFT
 $DIS_TYPE
  CONSTANT 2.0e5
#BOUNDARY_CONDITION
 $PCS_TYPE
  LIQUID_FLOW
 $PRIMARY_VARIABLE
  PRESSURE1
 $GEO_TYPE
  POLYLINE RIGHT
 $DIS_TYPE
  CONSTANT 1.0e5
#BOUNDARY_CONDITION
 $PCS_TYPE
  HEAT_TRANSPORT
 $PRIMARY_VARIABLE
  TEMPERATURE1
 $GEO_TYPE
  POLYLINE LEFT
 $DIS_TYPE
  CONSTANT 60
#STOP
""",
    ".st": """#SOURCE_TERM
 $PCS_TYPE
  LIQUID_FLOW
 $PRIMARY_VARIABLE
  PRESSURE1
 $GEO_TYPE
  POINT WELL
 $DIS_TYPE
  CONSTANT -1.0e-6
#STOP
""",
    ".ic": """#INITIAL_CONDITION
 $PCS_TYPE
  LIQUID_FLOW
 $PRIMARY_VARIABLE
  PRESSURE1
 $GEO_TYPE
  DOMAIN
 $DIS_TYPE
  CONSTANT 1.0e5
#INITIAL_CONDITION
 $PCS_TYPE
  HEAT_TRANSPORT
 $PRIMARY_VARIABLE
  TEMPERATURE1
 $GEO_TYPE
  DOMAIN
 $DIS_TYPE
  CONSTANT 10
#STOP
""",
    ".mmp": """#MEDIUM_PROPERTIES
 $GEOMETRY_DIMENSION
  2
 $GEOMETRY_AREA
  1.0
 $POROSITY
  1 0.2
 $STORAGE
  1 1.0e-9
 $PERMEABILITY_TENSOR
  ORTHOTROPIC 1.0e-13 2.0e-14
#MEDIUM_PROPERTIES
 $GEOMETRY_DIMENSION
  2
 $GEOMETRY_AREA
  1.0
 $POROSITY
  1 0.3
 $STORAGE
  1 2.0e-9
 $PERMEABILITY_TENSOR
  ISOTROPIC 5.0e-14
#STOP
""",
    ".msp": """#SOLID_PROPERTIES
 $DENSITY
  1 2500.0
 $THERMAL
  EXPANSION 1.0e-5
  CAPACITY
  1 800
  CONDUCTIVITY
  1 2.5
#SOLID_PROPERTIES
 $DENSITY
  1 2600.0
  CAPACITY
  1 900
  CONDUCTIVITY
  1 3.0
#STOP
""",
    ".mfp": """#FLUID_PROPERTIES
 $FLUID_TYPE
  LIQUID
 $DENSITY
  1 1000.0
 $VISCOSITY
  1 1.0e-3
 $SPECIFIC_HEAT_CAPACITY
  1 4200.0
 $HEAT_CONDUCTIVITY
  1 0.6
#STOP
""",
    ".out": """#OUTPUT
 $NOD_VALUES
  PRESSURE1
  TEMPERATURE1
 $GEO_TYPE
  POINT OBS
 $DAT_TYPE
  TECPLOT
#OUTPUT
 $PCS_TYPE
  HEAT_TRANSPORT
 $NOD_VALUES
  TEMPERATURE1
 $GEO_TYPE
  DOMAIN
 $DAT_TYPE
  TECPLOT
 $TIM_TYPE
  STEPS 1
#OUTPUT
 $PCS_TYPE
  LIQUID_FLOW
 $NOD_VALUES
  PRESSURE1
 $GEO_TYPE
  DOMAIN
 $DAT_TYPE
  VTK
 $TIM_TYPE
  3
#STOP
""",
}


def write_section_deck(directory, base="sec", flow="LIQUID_FLOW"):
    """Writes the deck; flow = "GROUNDWATER_FLOW" makes the flow process
    the HEAD form (primary variable HEAD, the tensor is the hydraulic
    conductivity, boundary heads in m).  Returns the deck's base path."""
    os.makedirs(directory, exist_ok=True)
    for ext, text in FILES.items():
        if flow == "GROUNDWATER_FLOW":
            text = (text.replace("LIQUID_FLOW", "GROUNDWATER_FLOW").replace("PRESSURE1", "HEAD")
                    .replace("CONSTANT 2.0e5", "CONSTANT 20.0")
                    .replace("CONSTANT 1.0e5", "CONSTANT 10.0")
                    .replace("ORTHOTROPIC 1.0e-13 2.0e-14", "ORTHOTROPIC 1.0e-6 2.0e-7")
                    .replace("ISOTROPIC 5.0e-14", "ISOTROPIC 5.0e-7"))
        with open(os.path.join(directory, base + ext), "w") as f:
            f.write(text)
    with open(os.path.join(directory, base + ".msh"), "w") as f:
        f.write(quad_xz_msh())
    return os.path.join(directory, base)

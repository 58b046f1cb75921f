"""Case directories in the reference's dictionary format.

A CaseSpec holds the parameters of one case; write_case() lays them out as the
reference's properties/{latticeProperties,macroProperties,simulation} and
start/boundaries files (examples/*/ shows the format), so that the same directory
drives both this library and the reference solver.  Key order inside each block keeps
the reference's first-substring-match lookups unambiguous (e.g. "g (..)" precedes any
other word containing a 'g' in the EOS block).
"""
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

LB_NAMES = {0: "D2Q9", 1: "D3Q15"}
BOX_NAMES = ["generic", "periodicX", "periodicY", "periodicXY", "periodicZ", "periodicXYZ",
             "periodicXYSpot"]
MODEL_NAMES = {0: "liMRT", 1: "myMRT", 2: "xuMRT", 3: "liTempMRT"}
EOS_NAMES = {0: "idealGas", 1: "Carnahan-Starling", 2: "Peng-Robinson", 3: "vanDerWaals"}


def patch_names(lbmodel, btype):
    """Boundary list names the box generator produces (latticeBox2D.c:258-261, latticeBox3D.c:259-327)."""
    if lbmodel == 0:
        return [] if btype == 3 else ["X0", "X1", "Y0", "Y1"]
    return {0: ["X0", "X1", "Y0", "Y1", "Z0", "Z1"], 3: ["Z0", "Z1"], 4: ["X0", "X1", "Y0", "Y1"],
            5: [], 6: ["Z0", "Z1", "Spot"]}[btype]


@dataclass
class BC:
    type: str = "periodic"
    value: Optional[Tuple[float, ...]] = None     # fixedT/fixedRandomT: (T,), fixedU: (ux, uy, uz)
    pert: Optional[float] = None                  # fixedRandomT percentage
    duration: Optional[int] = None


@dataclass
class CaseSpec:
    lbmodel: int = 0
    nx: int = 32
    ny: int = 32
    nz: int = 1
    btype: int = 3
    spot_radius: float = 0.0
    spot_centre: Tuple[float, float] = (0.0, 0.0)
    # EOS (defaults: the reference examples' van der Waals constants)
    eos: int = 3
    a: float = 0.5
    b: float = 4.0
    w_pr: float = 0.344
    R: float = 1.0
    G: float = -1.0
    M: float = 1.0
    Cv: float = 1.0
    rho0: float = 0.0
    g: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    fext: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    # write R, M, G, g, extForce, rho_0 and Cv before the model line: the reference's
    # first-substring lookups (dictIO/singleEntry.c:14) then find these keys and not the
    # "R" of "Peng-Robinson" or the "g" of "Carnahan-Starling"
    eos_keys_first: bool = False
    # f
    f_model: int = 0
    f_Lambda: Tuple[float, ...] = (1.0, 1.1, 1.1, 1.0, 1.1, 1.0, 1.1, 1.0, 1.0)
    f_sigma: float = 0.125
    f_st: int = 0
    f_kappa: float = 0.0
    # g (None disables the energy field)
    g_model: Optional[int] = 1
    g_Lambda: Tuple[float, ...] = (0.0, 1.0, 1.0, 1.2, 1.0, 1.2, 1.0, 1.0, 1.0)
    g_alpha1: float = 0.25
    g_alpha2: float = 0.25
    f_bc: Dict[str, BC] = field(default_factory=dict)
    g_bc: Dict[str, BC] = field(default_factory=dict)
    start: int = 0
    end: int = 1000000
    write_interval: int = 1000000
    data_format: str = "asciiRaw"

    def patches(self) -> List[str]:
        return patch_names(self.lbmodel, self.btype)


def _fmt(x):
    return repr(float(x))


def _vec(v):
    return "(" + " ".join(_fmt(x) for x in v) + ")"


def write_case(path, spec: CaseSpec):
    os.makedirs(os.path.join(path, "properties"), exist_ok=True)
    os.makedirs(os.path.join(path, "start"), exist_ok=True)
    lat = ["Nx %d;" % spec.nx, "", "Ny %d;" % spec.ny, ""]
    if spec.lbmodel == 1:
        lat += ["Nz %d;" % spec.nz, ""]
    lat += ["LBModel %s;" % LB_NAMES[spec.lbmodel], "", "boundaryType %s;" % BOX_NAMES[spec.btype]]
    if spec.btype == 6:
        lat += ["", "spot", "{", "    radius %s;" % _fmt(spec.spot_radius),
                "    centre %s;" % _vec(spec.spot_centre), "}"]
    _write(path, "properties/latticeProperties", lat)

    mp = ["f", "{", "  Collision %s;" % MODEL_NAMES[spec.f_model], "",
          "  Lambda %s;" % _vec(spec.f_Lambda), "", "  sigma %s;" % _fmt(spec.f_sigma), ""]
    st = {0: "liSurfTen", 2: "xuSurfTen"}.get(spec.f_model, "liSurfTen")
    mp += ["  SurfaceTension %s;" % (st if spec.f_st else "none")]
    if spec.f_st:
        mp += ["", "  kappa_st %s;" % _fmt(spec.f_kappa)]
    mp += ["}", ""]
    if spec.g_model is not None:
        mp += ["g", "{", "  Collision %s;" % MODEL_NAMES[spec.g_model], "",
               "  Lambda %s;" % _vec(spec.g_Lambda), ""]
        if spec.g_model == 1:
            mp += ["  alpha_1 %s;" % _fmt(spec.g_alpha1), "", "  alpha_2 %s;" % _fmt(spec.g_alpha2)]
        mp += ["}", ""]
    common = ["    R %s;" % _fmt(spec.R), "", "    M %s;" % _fmt(spec.M), "",
              "    G %s;" % _fmt(spec.G), "", "    g %s;" % _vec(spec.g), "",
              "    extForce %s;" % _vec(spec.fext), "", "    rho_0 %s;" % _fmt(spec.rho0), "",
              "    Cv %s;" % _fmt(spec.Cv)]
    mp += ["EOS", "{"] + (common + [""] if spec.eos_keys_first else [])
    mp += ["    model %s;" % EOS_NAMES[spec.eos], ""]
    mp += {3: ["    a_vdw %s;" % _fmt(spec.a), "", "    b_vdw %s;" % _fmt(spec.b)],
           1: ["    a_cs %s;" % _fmt(spec.a), "", "    b_cs %s;" % _fmt(spec.b)],
           2: ["    a_pr %s;" % _fmt(spec.a), "", "    b_pr %s;" % _fmt(spec.b), "",
               "    w_pr %s;" % _fmt(spec.w_pr)],
           0: []}[spec.eos]
    mp += ([] if spec.eos_keys_first else [""] + common) + ["}"]
    _write(path, "properties/macroProperties", mp)

    _write(path, "properties/simulation",
           ["startTime %d;" % spec.start, "endTime %d;" % spec.end,
            "writeInterval %d;" % spec.write_interval, "dataFormat %s;" % spec.data_format])

    bd = []
    fields = [("f", spec.f_bc)] + ([("g", spec.g_bc)] if spec.g_model is not None else [])
    for fname, bcs in fields:
        bd += [fname, "{", ""]
        for p in spec.patches():
            bc = bcs.get(p, BC())
            bd += ["    %s" % p, "    {", "        type %s;" % bc.type]
            if bc.value is not None:
                if bc.type == "fixedU":
                    bd += ["        value %s;" % _vec(bc.value)]
                else:
                    bd += ["        value %s;" % _fmt(bc.value[0])]
            if bc.pert is not None:
                bd += ["        pert %s;" % _fmt(bc.pert)]
            if bc.duration is not None:
                bd += ["        duration %d;" % bc.duration]
            bd += ["    }", ""]
        bd += ["}", "", ""]
    _write(path, "start/boundaries", bd)
    # the field lists of start/initialFields (read by readVTKInfo.c for pvtu output); the
    # initial values themselves are set through the library
    _write(path, "start/initialFields", ["scalarFields", "{", "  rho", "  T", "  p", "}", "", "vectorFields", "{",
                                         "  U", "}", "", "pdfFields", "{", "  f"] + (["  g"] if spec.g_model is not None else [])
           + ["}", ""])
    return path


def _write(root, rel, lines):
    with open(os.path.join(root, rel), "w") as fh:
        fh.write("\n".join(lines) + "\n")

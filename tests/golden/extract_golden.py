"""Lift the golden data held by the reference's own test scripts into JSON
fixtures (run in a container where the reference tree is mounted; the JSON
files are committed, so tests never read the reference).

  createspectra_tri.json  <- TestScripts/CreateSpectra.Quanty:188-260, 664-736
  eigensystem_f6.json     <- TestScripts/Eigensystem.Quanty:571, 669
  eigensystem_hubbard.json <- TestScripts/Eigensystem.Quanty:1119-1136
  eigensystem_mlft.json   <- TestScripts/Eigensystem.Quanty:763-770 (real d5 ligand-field cluster)
                             and :1032-1039 (the same with the field along y)
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = open(os.path.join(REF, "TestScripts/CreateSpectra.Quanty")).read()
    blk = src[src.index("G0Dn1 = {"):src.index("-- some reference frequencies")]
    tabs = {}
    for m in re.finditer(r"(G\d(?:Dn|Up)\d) = \{ \{([^}]*)\} ,\s*\{([^}]*)\}", blk):
        tabs[m.group(1)] = {"a": [float(x) for x in m.group(2).split(",")],
                            "b": [float(x) for x in m.group(3).split(",")]}
    out = {
        "source": "TestScripts/CreateSpectra.Quanty:188-260 (NBath=5, unrestricted) and :664-736 "
                  "(NBath=6, last site restricted empty)",
        "omega": [-1.1, -0.5463, 0.0, 4.75, 100.1],
        "gamma": 0.1,
        "threshold_block1": 1e-12,
        "threshold_block2": 1e-13,
        "model": "impurity (modes 0,1) + chain; H0 = U n0 n1 + t hopping, H1 = U n0 n1 + i t hopping, "
                 "H2 = i t hopping; U=3, t=1; TDn = c+_0, TUp = c+_1; psi1 = (|10..>+|01..>)/sqrt2, "
                 "psi2 = (|10..> - 2i|01..>)/sqrt5",
        "tables": tabs,
    }
    json.dump(out, open(os.path.join(HERE, "createspectra_tri.json"), "w"), indent=1)
    src = open(os.path.join(REF, "TestScripts/Eigensystem.Quanty")).read()

    def grab(key):
        i = src.index(key)
        return [float(x) for x in src[i:src.index("}", i)].split("{")[1].split(",")]

    json.dump({"source": "TestScripts/Eigensystem.Quanty:571 (real) and :669 (complex); energies from Mathematica",
               "model": "f shell, 14 modes (dn = even, up = odd), 6 electrons; H = U(F0=0,F2=10,F4=6.25,F6=4) "
                        "+ 0.1 ldots; complex case adds 1e-6 Ly + 2e-6 Sy",
               "real": grab("expectedEn = {-12.11458547046453"),
               "complex": grab("expectedEn = {-12.11458548596329")},
              open(os.path.join(HERE, "eigensystem_f6.json"), "w"), indent=1)
    json.dump({"source": "TestScripts/Eigensystem.Quanty:1119-1136 (energies from Quanty; the script "
                         "checks them to 1e-4)",
               "model": "1D Hubbard ring of 8 sites, modes 2i = site i down, 2i+1 = site i up; hopping "
                        "+1 between nearest and +0.1 between next-nearest neighbours (spin-diagonal, "
                        "periodic), U = 3.5 n_dn n_up per site; 4 down and 4 up electrons",
               "nsites": 8, "t1": 1.0, "t2": 0.1, "U": 3.5,
               "energies": grab("expectedEn = { -5.0089945483798")},
              open(os.path.join(HERE, "eigensystem_hubbard.json"), "w"), indent=1)
    i = src.index("::mlft::")
    i = src.index("valuesExpected = {", i)
    blk = src[i:src.index("NF=20", i)]
    rows = [[float(x) for x in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", blk)]
    i = src.index("Complex ligand field theory")
    i = src.index("valuesExpected = {", i)
    blk = src[i:src.index("Hamiltonian", i)]
    rows_y = [[float(x) for x in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", blk)]
    json.dump({"source": "TestScripts/Eigensystem.Quanty:763-770 (Quanty output; the script checks to 2e-6)",
               "model": "d5 ligand-field cluster, NF 20: 3d modes 0-9 and ligand d modes 10-19 (dn even, up "
                        "odd); U 5.5, Delta 8, F2 9.35, F4 5.78, 10Dq 0.67, 10DqL 1.06, Veg 1.92, Vt2g 1.15, "
                        "zeta 0.04, B 1e-6 (2 Sz + Lz on 3d); start: 5 in 3d, 10 in L; 3d occupation 5..7",
               "observables": ["H", "S^2", "L^2", "J^2", "Sz_3d", "Lz_3d", "ldots_3d", "F2_3d", "F4_3d",
                               "Neg_3d", "Nt2g_3d", "Neg_Ld", "Nt2g_Ld", "N_3d"],
               "values": rows,
               "source_y": "TestScripts/Eigensystem.Quanty:1032-1039: B (2 Sy + Ly) instead of B (2 Sz + Lz); "
                           "columns 5 and 6 are Sy_3d and Ly_3d",
               "values_y": rows_y},
              open(os.path.join(HERE, "eigensystem_mlft.json"), "w"), indent=1)


if __name__ == "__main__":
    main()

"""Writes the converter golden fixtures (run once, in the build container).

The FCIDUMP -> _IN converters' byte output (line order, which spin block receives a
permutation under the shared dedupe dictionary, coefficient text) is not fixed by any
mathematical check, so it is pinned against the reference's own converter
(util/convert_fcidump.py: conv_rhf with -r, conv_uhf with -u) run on small synthetic
FCIDUMPs (RHF: synth_fcidump; UHF: synth_uhf_integrals + write_fcidump_uhf).  Output: for
each case, <case>_fcidump (input) and <case>_IN (the reference converter's output).
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from sfci_amd import integrals as I  # noqa: E402

REF = "/root/reference/util/convert_fcidump.py"
CASES = [("g3o2e", 3, 2, 11, "-r"), ("g4o4e", 4, 4, 12, "-r"), ("u3o2e", 3, 2, 13, "-u"), ("u4o4e", 4, 4, 14, "-u")]


def main():
    for name, n_orb, n_e, seed, flag in CASES:
        d = tempfile.mkdtemp()
        fc = os.path.join(d, name + "_fcidump")
        if flag == "-r":
            I.synth_fcidump(fc, n_orb, n_e, seed)
        else:
            h_a, h_b, g_aa, g_bb, g_ab, ec = I.synth_uhf_integrals(n_orb, seed)
            I.write_fcidump_uhf(fc, h_a, h_b, g_aa, g_bb, g_ab, n_e, ec)
        subprocess.run([sys.executable, REF, flag, name + "_fcidump"], cwd=d, check=True,
                       stdout=subprocess.DEVNULL)
        shutil.copy(fc, os.path.join(HERE, name + "_fcidump"))
        shutil.copy(os.path.join(d, name + "_IN"), os.path.join(HERE, name + "_IN"))
        print("wrote", name)


if __name__ == "__main__":
    main()

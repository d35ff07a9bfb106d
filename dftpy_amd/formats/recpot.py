"""CASTEP reciprocal-space local pseudopotential (``.recpot``), the layout of
functional/pseudo/recpot.py:11-62 of the reference."""
import numpy as np

from ..constants import Units
from ._util import open_text


def read_recpot(path):
    """Returns (q in 1/bohr, v(q) in Ha bohr^3, zval)
    (functional/pseudo/recpot.py:11-42)."""
    with open_text(path) as f:
        lines = f.readlines()
    ibegin, iend = 0, None
    for i, line in enumerate(lines):
        if "END COMMENT" in line:
            ibegin = i + 3
        elif ibegin > 1 and (line.strip() == "1000" or len(line.strip()) == 1):
            iend = i
            break
    if iend is None:
        raise ValueError("recpot file %s has no terminating '1000' line" % path)
    gmax = float(lines[ibegin - 1].strip()) * Units.Bohr
    v = np.array(" ".join(l.strip() for l in lines[ibegin:iend]).split(), dtype=np.float64)
    v = v / Units.Ha / Units.Bohr ** 3
    q = np.linspace(0, gmax, num=len(v))
    zval = round((v[0] - v[1]) * (q[1] ** 2) / (4.0 * np.pi))
    return q, v, zval


def write_recpot(path, q, v, header="dftpy_amd"):
    """recpot writer in the layout of functional/pseudo/recpot.py:44-62."""
    gmax = q[-1] / Units.Bohr
    vv = np.asarray(v) * Units.Ha * Units.Bohr ** 3
    with open_text(path, "w") as fw:
        fw.write("START COMMENT\n" + header + "\nEND COMMENT\n3     5\n" + str(gmax) + "\n")
        n3 = vv.size // 3
        for row in vv[: n3 * 3].reshape(-1, 3):
            fw.write("     {0[0]:22.15E}     {0[1]:22.15E}     {0[2]:22.15E}\n".format(row))
        rest = vv[n3 * 3:]
        for x in rest:
            # separated: the reference's writer fuses negative leftovers
            fw.write("     {0:22.15E}".format(x))
        if rest.size:
            fw.write("\n")
        fw.write("1000\n")

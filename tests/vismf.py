"""An independent reader of AMReX's VisMF MultiFab files (version-1 text
header, FabOnDisk entries, FAB headers with the IEEE little-endian double
descriptor), for checking the files the engine writes."""
import os
import re

import numpy as np


def read_vismf(prefix):
    """(ncomp, ngrow, boxes, fabs) of a VisMF MultiFab, parsed from its text
    header and the FAB headers of its data file."""
    lines = open(prefix + "_H").read().split("\n")
    vers, how, ncomp, ngrow = int(lines[0]), int(lines[1]), int(lines[2]), int(lines[3])
    assert (vers, how) == (1, 0)
    m = re.match(r"\((\d+) 0", lines[4])
    nbox = int(m.group(1))
    boxre = re.compile(r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(0,0,0\)\)")
    boxes = [tuple(map(int, boxre.match(lines[5 + b]).groups())) for b in range(nbox)]
    assert lines[5 + nbox] == ")" and int(lines[6 + nbox]) == nbox
    fabs = []
    for b in range(nbox):
        tag, name, off = lines[7 + nbox + b].split()
        assert tag == "FabOnDisk:"
        with open(os.path.join(os.path.dirname(prefix), name), "rb") as fh:
            fh.seek(int(off))
            head = fh.readline().decode()
            hm = re.match(r"FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(8 7 6 5 4 3 2 1\)\)\)"
                          r"\(\((-?\d+),(-?\d+),(-?\d+)\) \((-?\d+),(-?\d+),(-?\d+)\) \(0,0,0\)\) (\d+)\n", head)
            lo, hi, nc = hm.groups()[:3], hm.groups()[3:6], int(hm.group(7))
            shape = tuple(int(h) - int(l) + 1 for l, h in zip(lo, hi))
            assert nc == ncomp
            assert tuple(int(l) + ngrow for l in lo) == boxes[b][:3]
            data = np.fromfile(fh, dtype="<f8", count=nc * int(np.prod(shape)))
            fabs.append(data.reshape((nc,) + shape[::-1]))
    return ncomp, ngrow, boxes, fabs

"""History output (.hst) of a device mesh in the reference's text format
(outputs/history.cpp:355-452): a descriptor line, a column header, then one line per
output with time, dt and the 11 MHD sums, each printed with the `data_format`
(default " %12.5e", outputs.cpp:320-321).  The sums come from `ath_gpu_history`."""
from __future__ import annotations

NAMES = ["time", "dt", "mass", "1-mom", "2-mom", "3-mom", "1-KE", "2-KE", "3-KE",
         "tot-E", "1-ME", "2-ME", "3-ME"]
# the header pads each "[n]=name" field as history.cpp does
_PAD = {"time": 5, "dt": 7, "mass": 5, "1-mom": 4, "2-mom": 4, "3-mom": 4, "1-KE": 5,
        "2-KE": 5, "3-KE": 5, "tot-E": 3, "1-ME": 4, "2-ME": 4, "3-ME": 4}


def header():
    cols = "".join("[%d]=%s%s" % (n + 1, name, " " * _PAD[name])
                   for n, name in enumerate(NAMES))
    return "# Athena++ history data\n# " + cols + "\n"


class HistoryWriter:
    """Appends history lines for a GpuMesh to `path` (the header on the first line)."""

    def __init__(self, path, data_format="%12.5e"):
        self.path = path
        self.fmt = " " + data_format
        self.count = 0

    def line(self, time, dt, sums):
        return "".join(self.fmt % v for v in [time, dt] + list(sums)) + "\n"

    def write(self, mesh):
        t, dt, _ = mesh.state()
        sums = mesh.history()
        with open(self.path, "a") as f:
            if self.count == 0:
                f.write(header())
            f.write(self.line(t, dt, sums))
        self.count += 1
        return sums


def read(path):
    """(names, rows) of a history file written in this format."""
    rows = []
    with open(path) as f:
        for ln in f:
            if ln.startswith("#"):
                continue
            rows.append([float(x) for x in ln.split()])
    return NAMES, rows

"""Cache of tuned kernel schedules (``Session.autotune`` results).

One JSON file maps "N|precision" to the selected candidate per pass, the
measured device time per loop iteration and, when the caller supplies the
candidate table (``Session.schedule_info``), the resolved schedule of each
chosen candidate: a library rebuilt with a different candidate table then
no longer matches, and the entry is ignored.  Location: ``$CSG_SCHEDULE_CACHE``,
else ``~/.cache/csg_amd/schedules.json``.  Nothing is read unless a caller asks
for ``schedule="cached"``: the default schedules are fixed, so results do not
depend on what a previous tuning run measured.
"""
from __future__ import annotations

import json
import os
from typing import Optional


def path() -> str:
    return os.environ.get("CSG_SCHEDULE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "csg_amd", "schedules.json"))


def _read() -> dict:
    try:
        with open(path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load(n: int, precision: str, info: Optional[list] = None) -> Optional[tuple]:
    """The cached selection for (n, precision), or None.  With ``info`` (the
    library's candidate table), an entry that recorded its schedules is
    returned only if the same indices still select the same schedules."""
    e = _read().get(f"{int(n)}|{precision}")
    if not e:
        return None
    sel = tuple(int(x) for x in e["sel"])
    if info is not None and "sched" in e:
        for q in range(4):
            if not (0 <= sel[q] < len(info[q])) or dict(info[q][sel[q]]) != e["sched"][q]:
                return None
    return sel


def save(n: int, precision: str, sel, us_per_step: float, info: Optional[list] = None) -> None:
    d = _read()
    entry = {"sel": [int(x) for x in sel], "us_per_step": float(us_per_step)}
    if info is not None:
        entry["sched"] = [dict(info[q][int(sel[q])]) for q in range(4)]
    d[f"{int(n)}|{precision}"] = entry
    p = path()
    os.makedirs(os.path.dirname(p), exist_ok=True)
    tmp = p + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def tune(n: int, precision: str = "fp64", dt: float = 0.01, rounds: int = 2, *, radius: Optional[float] = None,
         device: int = 0, store: bool = True):
    """Tune the schedules of an n x n field in `precision` on a particle of
    radius 0.3 n (or `radius`) and store the choice: returns (sel, us)."""
    from .core_shell import Session, initial_fields
    c0, phi = initial_fields(n, radius=0.3 * n if radius is None else radius, width=2.0)
    s = Session(n, precision=precision, device=device)
    try:
        s.set_fields(c0, phi)
        sel, us = s.autotune(dt, rounds)
        info = s.schedule_info()
    finally:
        s.close()
    if store:
        save(n, precision, sel, us, info)
    return sel, us

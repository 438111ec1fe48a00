"""``save``/``load`` of a ``SymmetricTensor`` as one ``.npz`` (reference
``serialize.py:209-346``), and the MPS checkpoint directory built on it (reference
``network/mps.py:567-676``) that lets a served solve resume.

The file holds a JSON header in the reference's schema -- format version, the legs (side,
dual flag, name, and the space as provider ``{kind, name}`` plus sector fields with
degeneracies) and, for providers whose block values depend on a
coefficient convention (SU(2), fZ2), a gauge fingerprint -- plus one member per block.
Nothing derived is written: block order and shapes are recomputed from the legs, and the
header's block count is checked against that recomputation.  A file written under a
different gauge, a newer format, a wrong block count or with foreign members is refused.
Blocks load as float64 on the requested device.
"""

from __future__ import annotations

import json
import os
import pathlib

import numpy as np
import torch

from tenet.space import IN, OUT, GradedSpace, Leg
from tenet.structure import TensorStructure
from tenet.symmetry import FZ2Sector, SU2, SU2Sector, Trivial, TrivialSector, U1, U1Sector, fZ2

__all__ = ["FORMAT_VERSION", "MPS_FORMAT_VERSION", "GAUGES", "save", "load", "save_mps", "load_mps"]

FORMAT_VERSION = 1
MPS_FORMAT_VERSION = 1

_PROVIDERS = {"Trivial": Trivial, "U1": U1, "SU2": SU2, "fZ2": fZ2}
_SECTOR_OF = {
    "Trivial": lambda args: TrivialSector(),
    "U1": lambda args: U1Sector(int(args[0])),
    "SU2": lambda args: SU2Sector(int(args[0])),
    "fZ2": lambda args: FZ2Sector(int(args[0])),
}
# the coefficient conventions block values are written in: SU(2) in the TensorKit
# SU2Irrep gauge (Condon-Shortley CG, descending m; F pinned to the reference fixtures),
# fZ2 with F = 1, fermionic R and the Koszul permutation sign
GAUGES = {
    "SU2": "3j=condon-shortley;cg=condon-shortley,m-descending;f=tks-su2irrep;r=(-1)^(a+b-c);fs=(-1)^2j",
    "fZ2": "f=1;r=fermionic;fs=+1;twist=(-1)^p;perm=koszul",
}
# the reference writes "su2=racah;<racah build fingerprint>" (symmetry/su2.py:55)
REFERENCE_SU2_GAUGE_PREFIX = "su2=racah;"
VERSION = "0.1.0"  # written as the header's "tenet" field, as the reference does


def _sector_args(a) -> list:
    if isinstance(a, U1Sector):
        return [a.charge]
    if isinstance(a, SU2Sector):
        return [a.two_j]
    if isinstance(a, FZ2Sector):
        return [a.parity]
    return []


def _encode_leg(leg: Leg, axis: int) -> dict:
    """The reference's leg record (``serialize.py:165-178``): side, dual flag, name, and
    the space as ``{provider: {kind, name}, sectors: [[sector fields], degeneracy]}``."""
    if not (leg.name is None or isinstance(leg.name, (str, int))):
        raise TypeError(f"public axis {axis}: leg name {leg.name!r} cannot be written to JSON; "
                        "only None, str and int names are supported")
    kind = leg.provider.name
    if kind not in _PROVIDERS:
        raise ValueError(f"cannot serialize provider {kind!r}; known kinds are {sorted(_PROVIDERS)}")
    return {"side": "out" if leg.side is OUT else "in", "dual": bool(leg.dual), "name": leg.name,
            "space": {"provider": {"kind": kind, "name": kind},
                      "sectors": [[_sector_args(a), int(m)] for a, m in leg.space.sectors]}}


def _decode_leg(d: dict) -> Leg:
    """A leg record in the reference's schema (``serialize.py:181-188``)."""
    prov = d["space"]["provider"]
    kind = prov["kind"]
    if kind not in _PROVIDERS:
        raise KeyError(f"unknown provider kind {kind!r}; known kinds are {sorted(_PROVIDERS)}")
    if prov.get("name", kind) != kind:
        raise ValueError(f"provider {kind!r} renamed to {prov['name']!r} is a different symmetry; "
                         "only the default-named providers are available here")
    space = GradedSpace.new(_PROVIDERS[kind],
                            {_SECTOR_OF[kind](args): int(m) for args, m in d["space"]["sectors"]})
    return Leg(space, OUT if d["side"] == "out" else IN, bool(d["dual"]), d.get("name"))


def _gauge_accepted(kind: str, gauge: str) -> bool:
    """The running library's own fingerprint, or -- for SU(2) -- the reference's racah-derived one:
    the coefficient tables here agree with racah's on every row of the reference's F and R
    fixtures (to 4.4e-16), so such files are numerically comparable."""
    running = GAUGES.get(kind)
    if running is None or gauge == running:
        return True
    return kind == "SU2" and gauge.startswith(REFERENCE_SU2_GAUGE_PREFIX)


def save(t, path: str | os.PathLike, *, compress: bool = False) -> None:
    """Write ``t`` to ``path``: a JSON header plus one array per block (host float64)."""
    header = {
        "format": FORMAT_VERSION,
        "tenet": VERSION,
        "legs": [_encode_leg(leg, i) for i, leg in enumerate(t.legs)],
        "gauges": {k: GAUGES[k] for k in {leg.provider.name for leg in t.legs} if k in GAUGES},
        "num_blocks": t.structure.num_blocks,
    }
    blocks = t.numpy_blocks()
    writer = np.savez_compressed if compress else np.savez
    writer(path, header=np.array(json.dumps(header)),
           **{f"b{i}": np.asarray(b, dtype=np.float64) for i, b in enumerate(blocks)})


def load(path: str | os.PathLike, *, device=None, dtype=torch.float64):
    """Read a tensor written by ``save``; the structure is exactly the saved one."""
    from tenet.tensor import SymmetricTensor, default_device

    with np.load(path, allow_pickle=False) as z:
        header = json.loads(str(z["header"].item()))
        if header["format"] > FORMAT_VERSION:
            raise ValueError(f"{path}: file format {header['format']} is newer than this library's "
                             f"format {FORMAT_VERSION}")
        for kind, gauge in header.get("gauges", {}).items():
            running = GAUGES.get(kind)
            if not _gauge_accepted(kind, gauge):
                raise ValueError(f"{path}: {kind} coefficient gauge on disk is {gauge!r}, but this library "
                                 f"runs {running!r}; the block coefficients are not comparable")
        structure = TensorStructure(tuple(_decode_leg(d) for d in header["legs"]))
        n = header["num_blocks"]
        if n != structure.num_blocks:
            raise ValueError(f"{path}: header says {n} blocks but the structure it describes has "
                             f"{structure.num_blocks}")
        expected = {"header", *(f"b{i}" for i in range(n))}
        extra = sorted(set(z.files) - expected)
        if extra:
            raise ValueError(f"{path}: unexpected members {extra} in the archive")
        missing = sorted(expected - set(z.files))
        if missing:
            raise ValueError(f"{path}: missing block members {missing}")
        blocks = [z[f"b{i}"] for i in range(n)]
    dev = torch.device(device) if device is not None else default_device()
    return SymmetricTensor(structure, blocks, device=dev, dtype=dtype)


def save_mps(psi, path: str | pathlib.Path) -> None:
    """Write an MPS checkpoint directory: ``NNN.npz`` per site plus ``mps.json``
    (``format``, ``n_sites``, ``center``).  A non-empty destination is refused before
    anything is written."""
    path = pathlib.Path(path)
    if path.exists() and any(path.iterdir()):
        raise FileExistsError(f"{path}: destination is not empty; delete it or choose another")
    path.mkdir(parents=True, exist_ok=True)
    for n, t in enumerate(psi.sites):
        save(t, path / f"{n:03d}.npz")
    meta = {"format": MPS_FORMAT_VERSION, "n_sites": len(psi), "center": psi.center}
    (path / "mps.json").write_text(json.dumps(meta))


def load_mps(path: str | pathlib.Path, *, device=None):
    """Read a directory written by ``save_mps``; checks the file set, the centre and that
    consecutive sites' bond spaces agree."""
    from tenet.network.mps import MPS

    path = pathlib.Path(path)
    meta_path = path / "mps.json"
    if not meta_path.is_file():
        raise ValueError(f"{meta_path}: not found; an MPS directory must contain mps.json")
    meta = json.loads(meta_path.read_text())
    if meta["format"] > MPS_FORMAT_VERSION:
        raise ValueError(f"{meta_path}: MPS directory format {meta['format']} is newer than this "
                         f"build's format {MPS_FORMAT_VERSION}")
    n_sites, center = meta["n_sites"], meta["center"]
    expected = {f"{n:03d}.npz" for n in range(n_sites)}
    found = {p.name for p in path.iterdir()} - {"mps.json"}
    if found != expected:
        raise ValueError(f"{path}: mps.json says n_sites={n_sites}, but the directory is missing "
                         f"{sorted(expected - found)} and has unexpected {sorted(found - expected)}")
    if center is not None and center not in range(n_sites):
        raise ValueError(f"{meta_path}: center {center} is neither null nor in range({n_sites})")
    sites = [load(path / f"{n:03d}.npz", device=device) for n in range(n_sites)]
    for n in range(n_sites - 1):
        if sites[n].legs[2].space != sites[n + 1].legs[0].space:
            raise ValueError(f"{path}: site {n}'s right bond space does not match site {n + 1}'s "
                             "left bond space; the directory is corrupt")
    return MPS(sites, center)

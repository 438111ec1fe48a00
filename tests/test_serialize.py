"""``save``/``load`` of tensors and the MPS checkpoint directory (reference
``tests/test_serialize.py`` and ``network/mps.py:567-676``): exact round trips for every
provider kind, and every refusal the format promises."""

import json

import numpy as np
import pytest
import torch


def _legs(kind):
    from tenet import IN, OUT, GradedSpace, Leg
    from tenet.symmetry import SU2, U1, FZ2Sector, SU2Sector, U1Sector, fZ2

    if kind == "u1":
        V = GradedSpace.new(U1, {U1Sector(-1): 2, U1Sector(0): 3, U1Sector(1): 1})
    elif kind == "su2":
        V = GradedSpace.new(SU2, {SU2Sector(0): 2, SU2Sector(1): 2, SU2Sector(2): 1})
    else:
        V = GradedSpace.new(fZ2, {FZ2Sector(0): 2, FZ2Sector(1): 3})
    return (Leg(V, OUT), Leg(V, OUT, True, "p"), Leg(V, IN), Leg(V, IN, False, 7))


@pytest.mark.parametrize("kind", ["u1", "su2", "fz2"])
def test_round_trip_is_exact(tmp_path, kind):
    from tenet import SymmetricTensor
    from tenet.serialize import load, save

    t = SymmetricTensor.random(_legs(kind), seed=4, device="cpu")
    save(t, tmp_path / "t.npz")
    back = load(tmp_path / "t.npz", device="cpu")
    assert back.structure == t.structure
    assert torch.equal(back.data, t.data)


def _rewrite(path, mutate):
    with np.load(path, allow_pickle=False) as z:
        members = {k: z[k] for k in z.files}
    header = json.loads(str(members["header"].item()))
    mutate(header, members)
    members["header"] = np.array(json.dumps(header))
    np.savez(path, **members)


def test_refuses_a_foreign_gauge_a_newer_format_and_a_wrong_block_count(tmp_path):
    from tenet import SymmetricTensor
    from tenet.serialize import load, save

    t = SymmetricTensor.random(_legs("su2"), seed=1, device="cpu")
    p = tmp_path / "t.npz"
    for mutate, msg in (
        (lambda h, m: h["gauges"].__setitem__("SU2", "some-other-gauge"), "gauge"),
        (lambda h, m: h.__setitem__("format", 99), "newer"),
        (lambda h, m: h.__setitem__("num_blocks", h["num_blocks"] + 1), "blocks"),
        (lambda h, m: m.__setitem__("extra", np.zeros(1)), "unexpected"),
    ):
        save(t, p)
        _rewrite(p, mutate)
        with pytest.raises(ValueError, match=msg):
            load(p, device="cpu")


def test_refuses_an_unserializable_leg_name(tmp_path):
    from tenet import IN, OUT, GradedSpace, Leg, SymmetricTensor
    from tenet.serialize import save
    from tenet.symmetry import U1, U1Sector

    V = GradedSpace.new(U1, {U1Sector(0): 1})
    t = SymmetricTensor.random((Leg(V, OUT, False, ("a", 1)), Leg(V, IN)), seed=0, device="cpu")
    with pytest.raises(TypeError, match="public axis 0"):
        save(t, tmp_path / "t.npz")


@pytest.mark.gpu
def test_mps_checkpoint_round_trip_and_refusals(tmp_path):
    from tenet.network.models import random_chain, su2_random_chain
    from tenet.serialize import load_mps, save_mps

    for psi in (random_chain(8, 12, seed=3, device="cuda"), su2_random_chain(8, 6, seed=3, device="cuda")):
        psi.center = 2
        d = tmp_path / psi[0].provider.name
        save_mps(psi, d)
        back = load_mps(d, device="cuda")
        assert back.center == 2 and len(back) == len(psi)
        for a, b in zip(back.sites, psi.sites):
            assert a.structure == b.structure and torch.equal(a.data, b.data)
        with pytest.raises(FileExistsError):
            save_mps(psi, d)
        (d / "003.npz").unlink()
        with pytest.raises(ValueError, match="missing"):
            load_mps(d, device="cuda")


def test_reads_a_header_in_the_reference_schema(tmp_path):
    """A file laid out as the reference writes it (``serialize.py:165-188,259-264``): legs
    ``{side, dual, name, space: {provider: {kind, name}, sectors}}``, a ``tenet`` version
    field and, for SU(2), the reference's racah gauge fingerprint."""
    from tenet import SymmetricTensor
    from tenet.serialize import load

    for kind, gauges in (("u1", {}), ("su2", {"SU2": "su2=racah;sun-authority=racah-0.1.1"})):
        t = SymmetricTensor.random(_legs(kind), seed=9, device="cpu")
        prov = "U1" if kind == "u1" else "SU2"
        legs = []
        for l in t.legs:
            secs = [[[a.charge if kind == "u1" else a.two_j], m] for a, m in l.space.sectors]
            legs.append({"side": l.side.value, "dual": l.dual, "name": l.name,
                         "space": {"provider": {"kind": prov, "name": prov}, "sectors": secs}})
        header = {"format": 1, "tenet": "0.1.0", "legs": legs, "gauges": gauges,
                  "num_blocks": t.structure.num_blocks}
        p = tmp_path / f"{kind}.npz"
        np.savez(p, header=np.array(json.dumps(header)),
                 **{f"b{i}": b for i, b in enumerate(t.numpy_blocks())})
        back = load(p, device="cpu")
        assert back.structure == t.structure
        assert torch.equal(back.data, t.data)


def test_writes_the_reference_schema(tmp_path):
    from tenet import SymmetricTensor
    from tenet.serialize import save

    t = SymmetricTensor.random(_legs("fz2"), seed=2, device="cpu")
    save(t, tmp_path / "t.npz")
    with np.load(tmp_path / "t.npz", allow_pickle=False) as z:
        header = json.loads(str(z["header"].item()))
    assert header["tenet"] and header["format"] == 1
    leg = header["legs"][1]
    assert set(leg) == {"side", "dual", "name", "space"}
    assert leg["space"]["provider"] == {"kind": "fZ2", "name": "fZ2"}
    assert leg["space"]["sectors"] == [[[0], 2], [[1], 3]]
    assert (leg["side"], leg["dual"], leg["name"]) == ("out", True, "p")
    assert header["gauges"] == {"fZ2": "f=1;r=fermionic;fs=+1;twist=(-1)^p;perm=koszul"}

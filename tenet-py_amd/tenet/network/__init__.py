"""Tensor-network layer: MPS/MPO, environments, the two-site H_eff and DMRG."""

from tenet.network.dmrg import DMRG_out, Sweep, dmrg_, lanczos, spectrum, sweep_
from tenet.network.env import Env, measure_mpo, ones
from tenet.network.heff import Heff2
from tenet.network.mps import MPO, MPS, from_dense, overlap
from tenet.network.terms import Site, local_op, spin_half

__all__ = ["DMRG_out", "Env", "Heff2", "MPO", "MPS", "Site", "Sweep", "dmrg_", "from_dense", "lanczos",
           "local_op", "measure_mpo", "ones", "overlap", "spectrum", "spin_half", "sweep_"]

"""Write tests/golden/golden.npz from the CPU oracle.

Config "norb=8, nelec=(4,4)" (dim 70 x 70): state = random_state_vector(4900,
seed=0); U = Haar random_unitary(8, seed=1234); J = symmetric with upper
triangle U(-1, 1) drawn with seed 5678 (ffsim's _random_symmetric_matrix_uniform
convention, scale 2); lambda = U(-1, 1)^8 with seed 91011; t = 0.7.
Expected outputs come from the oracle's restatement of ffsim's schedules;
tests/test_oracle_*.py pin that oracle against closed forms.
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import oracle as O  # noqa: E402


def main() -> None:
    norb, nelec = 8, (4, 4)
    dim = math.comb(norb, 4) ** 2
    vec = O.random_state_vector(dim, np.random.default_rng(0))
    U = O.random_unitary(norb, np.random.default_rng(1234))
    J = O.random_symmetric_uniform(norb, np.random.default_rng(5678), scale=2.0)
    lam = np.random.default_rng(91011).uniform(-1, 1, size=norb)
    t = 0.7
    rotated = O.apply_orbital_rotation(vec, U, norb, nelec)
    dc = O.apply_diag_coulomb_evolution(rotated, J, t, norb, nelec)
    dcz = O.apply_diag_coulomb_evolution(rotated, J, t, norb, nelec, z_representation=True)
    numop = O.apply_num_op_sum_evolution(vec, lam, t, norb, nelec)
    contract = O.contract_diag_coulomb(vec, J, norb, nelec)
    contract_z = O.contract_diag_coulomb(vec, J, norb, nelec, z_representation=True)
    contract_num = O.contract_num_op_sum(vec, lam, norb, nelec)

    fixture = np.load(HERE / "orbital_rotation-0.npy")
    hf = np.zeros(math.comb(8, 5) ** 2, dtype=complex)
    hf[0] = 1
    special = O.orbital_rotation_closed_form(hf, fixture, fixture, 8, (5, 5))

    tables = {}
    for n in range(1, 9):
        for k in range(0, n + 1):
            tables[f"strings_{n}_{k}"] = O.cached_strings(n, k)
            tables[f"occslst_{n}_{k}"] = O.cached_occslst(n, k).astype(np.int64)
    for k in range(1, 8):
        for i in range(7):
            for pair in ((i, i + 1), (i + 1, i)):
                tables[f"zero_one_8_{k}_{pair[0]}_{pair[1]}"] = O.zero_one_subspace_indices(
                    8, k, pair
                ).astype(np.int64)

    np.savez_compressed(
        HERE / "golden.npz",
        norb=norb,
        nelec=np.array(nelec),
        vec=vec,
        U=U,
        J=J,
        lam=lam,
        time=t,
        rotated=rotated,
        diag_coulomb=dc,
        diag_coulomb_z=dcz,
        num_op=numop,
        contract=contract,
        contract_z=contract_z,
        contract_num=contract_num,
        special_case_expected=special,
        **tables,
    )
    print("wrote", HERE / "golden.npz")


if __name__ == "__main__":
    main()

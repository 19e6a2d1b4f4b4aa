"""Optional launch modes of the force call against the oracle: split-major wave order
(the four waves of a work-group share a j-range; the in-LDS pre-reduction of the
partial sums is off then), deferred neighbour-id stores, and both together.  Forces
within the precision's tolerance, nearest neighbours and lists bit-exact."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
import sapporo2_amd as sa
from _util import check_forces, check_nn_lists, load_engine, plummer_system

pytestmark = pytest.mark.gpu
TOL = {sa.DOUBLESINGLE: 1e-6, sa.DOUBLE: 1e-12}


@pytest.mark.parametrize("prec", [sa.DOUBLESINGLE, sa.DOUBLE])
def test_split_major_and_pipelined_lists(prec):
    n = 8192
    m, x, v, ids, h2 = plummer_system(n, seed=81)
    eps2 = 1e-4
    rng = np.random.Generator(np.random.PCG64(81))
    with sa.Engine(sa.FOURTH, prec) as eng:
        load_engine(eng, m, x, v, ids)
        eng.set_time(0.0)
        for ni in (100, 2000):
            sel = np.sort(rng.choice(n, ni, replace=False))
            ref = oracle.force(1, x[sel], v[sel], ids[sel], x, v, m, ids, eps2, h2=h2[sel])
            for smajor, pipe in ((1, 0), (0, 1), (1, 1)):
                print(f"modes prec={prec} ni={ni} split-major={smajor} pipeline={pipe}", flush=True)
                sys.stdout.flush()
                eng.set_split_major(bool(smajor))
                eng.set_ngb_pipeline(bool(pipe))
                eng.start_grav_calc(n, ids[sel], x[sel], v[sel], eps2, h2[sel])
                res = eng.get_grav_results(n, ni)
                label = f"ni={ni} smajor={smajor} pipe={pipe}"
                check_forces(res, ref, 1, TOL[prec], label=label)
                check_nn_lists(eng, res, prec == sa.DOUBLE, x[sel], ids[sel], x, ids, h2[sel], label=label)
        eng.set_split_major(False)
        eng.set_ngb_pipeline(False)

"""Contexts created and destroyed back to back, on several threads at once,
take the HIP streams and rocBLAS handles of destroyed contexts from
per-device pools (hfk_host.hip, hfk_scf.hip).  Reused streams and handles
must not change any result: every SCF here equals the one of a context that
ran alone."""
import threading

import numpy as np
import pytest

import hfk
from hfk import systems

pytestmark = pytest.mark.gpu


def _scf(ctx):
    r = ctx.scfHF(converge=1e-8, max_steps=60, diis_start=float("inf"), diis_size=6)
    return r.getEnergy, np.asarray(r.getDensity, dtype=np.float64).copy(), r.steps


def test_pooled_streams_and_handles_keep_results():
    mol = systems.water("6-31g")
    with hfk.Context(mol) as ctx:
        e_ref, d_ref, steps_ref = _scf(ctx)
    out, errors = [], []
    lock = threading.Lock()

    def work():
        try:
            for _ in range(5):
                with hfk.Context(mol) as c:
                    res = _scf(c)
                with lock:
                    out.append(res)
        except Exception as exc:   # reported by the main thread
            with lock:
                errors.append(repr(exc))

    threads = [threading.Thread(target=work) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(out) == 20
    for e, d, steps in out:
        assert steps == steps_ref
        assert abs(e - e_ref) < 1e-10
        assert np.max(np.abs(d - d_ref)) < 1e-10

"""Serving loop on the GPU: a job's result is the same alone and in a mix.

Each job's final mapping is checked against the oracle's score; the same job
run alone through a one-slot server, in a mix of jobs of uneven n on a
four-slot server, and directly through bm_refine_run must end with the same
mapping, score, step and move counts.
"""
import threading

import numpy as np
import pytest

import bmatch
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def jobs(count, seed=100):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(120, 700))
        inst = bmatch.Instance.powerlaw(seed + i, n, 12 * n, shuffle_frac=0.5)
        prm = bmatch.refine_params(pairs=1, triples=512 * (1 + i % 3), sets=64 * (i % 2), set_k=4 + i % 3,
                                   rounds=2, steps_per_check=3, seed=seed + 7 * i)
        out.append((inst, prm))
    return out


def direct(inst, prm, move_budget, patience):
    with bmatch.Engine(0) as e:
        e.upload(inst.n, inst.g1, inst.g2)
        e.set_mapping(inst.p0)
        e.refine_configure(prm)
        st = e.refine_run(move_budget=move_budget, patience=patience)
        return st, e.mapping()


def test_result_same_alone_and_in_mix():
    js = jobs(6)
    budget, patience = 0, 3
    alone = []
    with bmatch.Server(max_active=1, prep_threads=1) as srv:
        for inst, prm in js[:2]:
            jid = srv.submit(inst.n, inst.g1, inst.g2, inst.p0, prm, move_budget=budget, patience=patience)
            alone.append(srv.wait(jid, inst.n))
    with bmatch.Server(max_active=4, prep_threads=2) as srv:
        ids = [srv.submit(inst.n, inst.g1, inst.g2, inst.p0, prm, move_budget=budget, patience=patience)
               for inst, prm in js]
        mixed = [srv.wait(i, inst.n) for i, (inst, _) in zip(ids, js)]
        st = srv.stats()
    assert st["done"] == 6 and st["failed"] == 0 and st["peak_active"] >= 2
    for (inst, prm), (res, mp) in zip(js, mixed):
        assert res["status"] == 0, res["error"]
        assert res["final_score"] == ob.Oracle(inst.n, inst.g1, inst.g2).score_opt(mp)
        assert res["final_score"] > res["initial_score"]
        assert res["t_submit"] <= res["t_admit"] <= res["t_start"] <= res["t_done"]
    for (inst, prm), (ra, ma), (rm, mm) in zip(js, alone, mixed):
        np.testing.assert_array_equal(ma, mm)
        for f in ("initial_score", "final_score", "steps", "moves"):
            assert ra[f] == rm[f], f
        st, md = direct(inst, prm, budget, patience)
        np.testing.assert_array_equal(md, mm)
        assert (st["final_score"], st["steps"], st["moves"]) == (rm["final_score"], rm["steps"], rm["moves"])


def test_arrivals_while_running_and_budget_rule():
    js = jobs(5, seed=300)
    with bmatch.Server(max_active=3, prep_threads=2) as srv:
        got = {}

        def client():
            for k, (inst, prm) in enumerate(js):
                got[k] = srv.submit(inst.n, inst.g1, inst.g2, inst.p0, prm, move_budget=40, patience=3)

        t = threading.Thread(target=client)
        t.start()
        t.join()
        res = [srv.wait(got[k], inst.n) for k, (inst, _) in enumerate(js)]
    for (inst, prm), (r, mp) in zip(js, res):
        assert r["status"] == 0, r["error"]
        st, md = direct(inst, prm, 40, 3)
        np.testing.assert_array_equal(md, mp)
        assert st["moves"] == r["moves"]


def test_admission_against_the_budget():
    js = jobs(4, seed=500)
    biggest = max(bmatch.job_footprint(inst.n, len(inst.g1[0]), len(inst.g2[0]), prm) for inst, prm in js)
    with bmatch.Server(max_active=4, prep_threads=2, hbm_budget=biggest + (1 << 20)) as srv:
        ids = [srv.submit(inst.n, inst.g1, inst.g2, inst.p0, prm, max_steps=6, patience=2) for inst, prm in js]
        for i, (inst, _) in zip(ids, js):
            r, _ = srv.wait(i, inst.n)
            assert r["status"] == 0
        st = srv.stats()
        assert st["peak_active"] == 1 and st["peak_reserved_bytes"] <= biggest + (1 << 20)
        inst, prm = js[0]
        with pytest.raises(bmatch.BMError):
            # a job larger than the whole budget is refused at submission
            big = bmatch.Instance.powerlaw(9, 3000, 30000)
            srv.submit(big.n, big.g1, big.g2, big.p0, prm)


def test_failing_job_does_not_disturb_others():
    js = jobs(3, seed=700)
    with bmatch.Server(max_active=2, prep_threads=2) as srv:
        bad = js[0][0].p0.copy()
        bad[1] = js[0][0].n + 5  # outside [0, n]: bm_set_mapping refuses it
        ids = [srv.submit(js[0][0].n, js[0][0].g1, js[0][0].g2, bad, js[0][1])]
        ids += [srv.submit(inst.n, inst.g1, inst.g2, inst.p0, prm, patience=2) for inst, prm in js[1:]]
        r0, m0 = srv.wait(ids[0], js[0][0].n)
        assert r0["status"] == bmatch.BM_EINVAL and m0 is None
        for i, (inst, _) in zip(ids[1:], js[1:]):
            r, mp = srv.wait(i, inst.n)
            assert r["status"] == 0
        assert srv.stats()["failed"] == 1

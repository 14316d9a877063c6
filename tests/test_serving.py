"""The served request loop: requests arriving while others run are admitted between step
chunks (by HBM budget), their work overlaps on the device, and every result equals the
same request run alone."""
from __future__ import annotations

import numpy as np
import pytest

from paraphysicell_amd import abi, synth
from paraphysicell_amd.serving import Request, Server, footprint_bytes, run_alone

pytestmark = pytest.mark.gpu


def _requests(sizes, gaps, steps=23):
    out = []
    t = 0.0
    for k, (n, gap) in enumerate(zip(sizes, gaps)):
        t += gap
        dims = 2 + k % 2
        mesh, st, params = synth.ball_population(dims, n, 40.0 * (n / 100) ** (1 / dims), 400, 100 + k, 4, defs=2,
                                                 subs=1, springs=True)
        st["attachment_rate"] = 0.5
        st["detachment_rate"] = 0.05
        st.view("attack_rates")[:, 0] = 0.05
        st["cell_definition_index"] = np.arange(n) % 2
        out.append(Request(k, t, mesh, st, params, steps, seed=7 + k))
    return out


@pytest.mark.parametrize("budget_members", [2, 100])
def test_served_results_equal_alone(budget_members):
    print(f"serving, budget for {budget_members} members", flush=True)
    sizes = [300, 1200, 150, 800, 2000, 500]
    gaps = [0.0, 0.002, 0.0, 0.004, 0.001, 0.003]
    alone = _requests(sizes, gaps)
    run_alone(alone)
    served = _requests(sizes, gaps)
    one = max(footprint_bytes(q.mesh, 4, q.state.defs, q.state.subs, q.agents) for q in served)
    server = Server(budget_bytes=budget_members * one, chunk=5)
    done = server.run(served, realtime=True)
    server.close()
    assert len(done) == len(sizes)
    by_key = {q.key: q for q in done}
    for ref in alone:
        got = by_key[ref.key].result
        assert got["count"] == ref.result["count"], ref.key
        assert got["deaths"] == ref.result["deaths"], ref.key
        assert np.array_equal(got["positions"], ref.result["positions"]), ref.key
        assert np.array_equal(got["velocities"], ref.result["velocities"]), ref.key
    lat = [q.latency for q in done]
    assert all(x > 0 for x in lat)

"""Config 4 at its real sizes through the request server (GPU): a Poisson stream of single-sun requests over six
mixed canopies (square patches, triangles and triangle-fan disks; 31,564 to 1,838,932 primitives; 1-4 bands
each) arriving while earlier requests run.  Every served result must equal the solve of that request alone
(rt_run_band, 1e-5 relative, floor 1 W/m^2: the same rays, summed in another grouping), and the smallest model's
served results must match the FP64 oracle at the parity bound (1e-4)."""
import time

import numpy as np
import pytest

from helios_amd import native
from helios_amd import workloads as W
from oracle.oracle import OracleScene
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(ctx):
    models = W.serve_models(seed=4, nscenes=6)
    scs = [ctx.scene(s.transforms, s.types, s.twosided) for s, _ in models]
    reqs = W.serve_requests(models, 32, seed=5)
    srv = native.Server(ctx, slots=4, max_batch_suns=32)
    ids = [srv.add_model(g, req) for g, (_, req) in zip(scs, models)]
    rng = np.random.default_rng(6)
    tickets = []
    for (m, d, f, q, seed) in reqs:
        tickets.append(srv.submit(ids[m], d, f, q, seed))
        time.sleep(float(rng.exponential(0.05)))  # arrivals while earlier requests are on the device
    out = [srv.wait(t)[0] for t in tickets]
    st = srv.stats()
    srv.close()
    return models, scs, reqs, out, st


def test_config4_stream_served_equals_alone(served):
    models, scs, reqs, out, st = served
    assert max(s.P for s, _ in models) > 1_000_000 and min(s.P for s, _ in models) < 50_000
    assert len({r.B for _, r in models}) >= 3  # models of different band counts share the server
    big = [k for k, (m, *_r) in enumerate(reqs) if models[m][0].P > 1_000_000]
    assert len(big) >= 4  # the stream carries requests for the largest models
    worst = 0.0
    for k, (m, d, f, q, seed) in enumerate(reqs):
        alone = native.run_band(scs[m], W.single_sun_request(models[m][1], d, f, q, seed))
        worst = max(worst, float(rel_err(out[k], alone, floor=1.0).max()))
    assert worst <= 1e-5, worst
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] < len(reqs) and st["max_batch"] > 1, st  # continuous batching happened
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_config4_small_model_against_oracle(served):
    models, scs, reqs, out, st = served
    small = min(range(len(models)), key=lambda m: models[m][0].P)
    s, base = models[small]
    ks = [k for k, (m, *_r) in enumerate(reqs) if m == small]
    assert ks, "the stream has no request for the smallest model"
    m, d, f, q, seed = reqs[ks[0]]
    Ro = OracleScene.from_arrays(s).run_band(W.single_sun_request(base, d, f, q, seed))
    assert rel_err(out[ks[0]], Ro, floor=1.0).max() <= 1e-4

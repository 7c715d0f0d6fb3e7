"""Error paths on the GPU: an allocation the device-memory cap refuses is
reported as OVG_ENOMEM by the call that needed it (never OVG_OK with stale
results), and the context works again once the cap is lifted."""
import sys

import numpy as np
import pytest

import openvort_amd as ov
from openvort_amd import configs
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture
def capped():
    prev = ov.set_device_memory_limit(0)
    yield
    ov.set_device_memory_limit(prev)


def test_direct_sum_out_of_memory(capped):
    """ovg_velocity_at's direct sum needs a chunk-partials buffer; on a
    context that has only run the Barnes-Hut list traversal that buffer does
    not exist yet."""
    t, p, _ = configs.make("c2", rings=20, traverse=ov.TRAVERSE_LIST,
                           precision=ov.PREC_MIXED)
    slots = t.live[::7]
    say(f"velocity_at under a device-memory cap: {len(slots)} of {len(t.live)} points")
    with ov.Engine(p) as e:
        e.upload(t)
        e.update_tangle(0.0)
        e.synchronize()
        ov.set_device_memory_limit(ov.device_memory_in_use() + 4096)
        with pytest.raises(ov.OvgError) as ei:
            e.velocity_at(slots, 0.0)
        assert ei.value.code == ov.ENOMEM
        assert "memory limit" in str(ei.value)
        ov.set_device_memory_limit(0)
        vs, vels = e.velocity_at(slots, 0.0)
        # the same call on a fresh context without a cap
        with ov.Engine(p) as f:
            f.upload(t)
            f.update_tangle(0.0)
            vs2, vels2 = f.velocity_at(slots, 0.0)
    assert np.array_equal(vs, vs2) and np.array_equal(vels, vels2)
    assert np.abs(vels).max() > 0


def test_update_velocities_out_of_memory(capped):
    """The direct-method evaluation reports a refused buffer instead of
    returning the velocities of an earlier call."""
    t, p, _ = configs.make("c1", rings=6)
    say(f"direct update_tangle under a cap: {len(t.live)} points")
    with ov.Engine(p) as e:
        e.upload(t)
        e.synchronize()
        ov.set_device_memory_limit(ov.device_memory_in_use() + 4096)
        with pytest.raises(ov.OvgError) as ei:
            e.update_tangle(0.0)
        assert ei.value.code == ov.ENOMEM
        ov.set_device_memory_limit(0)
        e.update_tangle(0.0)
        o = e.download()
    assert np.abs(o["vels"][t.live]).max() > 0


def test_tree_build_out_of_memory(capped):
    t, p, _ = configs.make("c2", rings=20, traverse=ov.TRAVERSE_LIST,
                           precision=ov.PREC_MIXED)
    say("Barnes-Hut update_tangle under a cap")
    with ov.Engine(p) as e:
        e.upload(t)
        e.synchronize()
        ov.set_device_memory_limit(ov.device_memory_in_use() + 4096)
        with pytest.raises(ov.OvgError) as ei:
            e.update_tangle(0.0)
        assert ei.value.code == ov.ENOMEM
        ov.set_device_memory_limit(0)
        e.update_tangle(0.0)
        a = e.download()["vels"]
    with ov.Engine(p) as f:
        f.upload(t)
        f.update_tangle(0.0)
        b = f.download()["vels"]
    live = t.live
    assert rel_l2(a[live], b[live]) == 0.0

"""Networks above 2^32 synapses on one GPU: 4.5e7 LIF neurons with in-degree 100 (4.5e9
synapses, 64-bit CSR slots, grid-stride kernels over more than 2^32 work items, the 64-bit E->E
transpose of pair STDP).  Checked by size-independent properties: the row pointers are
monotone and end at N x k_in; every spike delivers its row: events = sum over neurons of spike
count x out-degree; the run spikes."""
import numpy as np
import pytest

from neuralsim_amd import configs

pytestmark = pytest.mark.gpu


def test_over_2_32_synapses_properties(nsx):
    n = 45_000_000
    spec = configs.lif_large_network(n=n, drive_scale=60.0)
    assert spec.m > (1 << 32)
    print(f"case large {spec.name} m={spec.m}", flush=True)
    with nsx.SimulationEngine.from_spec(spec) as eng:
        eng.set_graph_steps(25)
        st = eng.step(100)
        rp = eng.row_ptr()
        cnt = eng.read_state(["spike_count"])["spike_count"].astype(np.uint64)
    assert int(rp[0]) == 0 and int(rp[-1]) == spec.m
    outdeg = np.diff(rp.astype(np.int64))
    assert outdeg.min() >= 0
    assert st.total_spikes == int(cnt.sum()) > 0
    assert st.total_events == int((cnt * outdeg.astype(np.uint64)).sum())

"""swl_engine2d_diagnostics_noalloc (GPU): the engine diagnostics with the inspection's partials in the
library's fixed device scratch array instead of a per-call allocation.  It must return exactly what
swl_engine2d_diagnostics returns, on consecutive calls whose second one covers a smaller grid (the scratch
then still holds the first call's partials beyond the second grid's tiles, which must not be read)."""
import pytest

from cpu_sessions import CpuSession2D
from seiswavelab_curvilinear_amd import Precision, workloads
from seiswavelab_curvilinear_amd.engine import Engine2D, mesh_from_arrays, source_scale
from test_oracle_2d import ref_case

pytestmark = pytest.mark.gpu


def stepped_engine(p, precision, steps):
    p.precision = Precision.Float64
    o = CpuSession2D("orc", p)
    e = Engine2D(mesh_from_arrays(o.model, o.layout(), p, precision=precision), o.probe_indices(), 0)
    for n in range(steps):
        e.step(n * o.dt, o.dt, p.frequency, source_scale(p), n)
    return e


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_diagnostics_equal_allocating_diagnostics(precision):
    big = workloads.grf_terrain_5001x2001(seed=11, nx=400, nz=150, steps=60, topo_rms=6.0).params  # 460x180 nodes
    small = ref_case("sponge_tilted_force")  # 70x60 nodes
    print(f"[no-alloc] precision {precision}", flush=True)
    for p, steps in ((big, 60), (small, 40)):
        e = stepped_engine(p, precision, steps)
        want = e.diagnostics()
        got = e.diagnostics_noalloc()
        assert want["energy"] > 0
        assert got == want
        assert e.diagnostics_noalloc() == want  # a repeated call reads the same
        e.close()

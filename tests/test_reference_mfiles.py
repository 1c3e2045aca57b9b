"""The reference's own model files run unchanged: models/schnakenberg.m,
brusselator.m, allencahn.m and test/models/linear.m, logistic.m, read where
they lie in the reference tree, compile through this front end
(tsphere.mfront, standing in for numbl) to the same IR, op lists and
generated HIP source as the package's copies under turing-sphere_amd/mfiles/
(which word their header comments differently).  The source text is only
parsed by this package's own front end; nothing of the reference runs.
Skipped where the reference tree is absent."""
import dataclasses
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFILES = os.path.join(ROOT, "turing-sphere_amd", "mfiles")
REF = "/root/reference"

MODELS = {  # key -> (path in the reference, params, state, view)
    "schnakenberg": ("models/schnakenberg.m", ["a", "b", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "brusselator": ("models/brusselator.m", ["A", "B", "D1", "D2", "dt"], ["U", "V"], ["u", "v"]),
    "allencahn": ("models/allencahn.m", ["eps2", "dt"], ["U"], ["u"]),
    "linear": ("test/models/linear.m", ["c", "D", "dt"], ["U"], ["u"]),
    "logistic": ("test/models/logistic.m", ["r", "D", "dt"], ["U"], ["u"]),
}


def _strip(x):
    """An IR tree as nested tuples without source positions (the two files'
    comments differ, so their offsets do)."""
    if dataclasses.is_dataclass(x):
        return (type(x).__name__,) + tuple(
            (f.name, _strip(getattr(x, f.name))) for f in dataclasses.fields(x)
            if f.name != "span")
    if isinstance(x, (list, tuple)):
        return tuple(_strip(v) for v in x)
    return x


def _compile(src, params, state, view, nlm=2080, npts=8192):
    from tsphere.mfront import CompiledModel, model_bindings
    from tsphere.plan import HostBuffers, ModelPlan, PlanSpec

    cm = CompiledModel(src, model_bindings(state, params, npts, nlm),
                       {"npts": npts, "nlm": nlm})
    n = len(state) + len(view)
    fns = [cm.specialize("init", n), cm.specialize("step", n)]
    cm.finish()
    fb = list(state) + [None] * len(view)
    out = {"ir": [_strip(f.body) for f in fns]}
    for batch in (False, True):
        host = HostBuffers(allocate=False)
        host.ensure_group(state, 2 * nlm)
        plans = [ModelPlan.create(None, PlanSpec(f, fb), host, batch, lbl)
                 for f, lbl in zip(fns, ("init", "step"))]
        out[f"ops{int(batch)}"] = [p.describe() for p in plans]
        out[f"src{int(batch)}"] = [p.kernel_source() for p in plans]
    return out


@pytest.mark.parametrize("key", sorted(MODELS))
def test_reference_model_file_compiles_as_package_copy(key):
    path, params, state, view = MODELS[key]
    ref_path = os.path.join(REF, path)
    if not os.path.exists(ref_path):
        pytest.skip("reference tree absent")
    with open(ref_path) as f:
        theirs = _compile(f.read(), params, state, view)
    with open(os.path.join(MFILES, key + ".m")) as f:
        ours = _compile(f.read(), params, state, view)
    assert theirs["ir"] and all(theirs["ir"])
    for k in ours:
        assert theirs[k] == ours[k], k

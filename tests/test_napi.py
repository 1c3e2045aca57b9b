"""The N-API binding (turing-sphere_amd/napi) that a node host -- the
reference's TypeScript ShtPlan / ModelPlan -- would load: it loads in node,
exports the calls, and reports the reference's configuration errors with its
messages (layout.ts:34-43).  CPU only; the GPU test drives a transform."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "turing-sphere_amd", "napi", "tsphere_napi.node")
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON),
                                reason="node or the built addon is not available")


def node(js: str) -> dict:
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_addon_loads_and_exports():
    from tsphere import _lib

    out = node(f"""const a = require({json.dumps(ADDON)});
      console.log(JSON.stringify({{version: a.version(), names: Object.getOwnPropertyNames(a).sort()}}));""")
    assert out["version"] == _lib.load().tsphere_version()
    for name in ("planCreate", "planQuery", "planDestroy", "synth", "analys", "malloc", "free",
                 "upload", "download", "moduleCompile", "moduleKernel", "kernelLaunch",
                 "graphBegin", "graphEnd", "graphLaunch", "streamCreate", "streamSync",
                 "legSynth", "legAnalys", "fourierTable", "kernelLaunchGrid", "legImexSynth",
                 "legSpectralImex"):
        assert name in out["names"]


def test_addon_config_errors():
    out = node(f"""const a = require({json.dumps(ADDON)});
      const r = [];
      for (const cfg of [{{lmax: 31, mmax: 31, nlat: 31, nphi: 64}},
                         {{lmax: 31, mmax: 31, nlat: 32, nphi: 62}}]) {{
        try {{ a.planCreate(cfg); r.push('created'); }}
        catch (e) {{ r.push([e.name, e.message, e.code]); }}
      }}
      let te = null;
      try {{ a.synth(1, 2, 3); }} catch (e) {{ te = e.name; }}
      let le = null;
      try {{ a.planCreate({{lmax: 31, mmax: 31, nlat: 32, nphi: 64, legendre: 'fp16'}}); }}
      catch (e) {{ le = [e.name, e.message]; }}
      console.log(JSON.stringify({{r, te, le}}));""")
    (n1, m1, c1), (n2, m2, c2) = out["r"]
    assert n1 == n2 == "ShtConfigError" and c1 == c2 == 1
    assert "nlat must be an integer > lmax" in m1 and "nphi must be an integer >= 2*mmax+1" in m2
    assert out["te"] == "TypeError"   # plain numbers are not handles
    # an unknown Legendre form is refused, not mapped to FP32
    assert out["le"][0] == "TypeError" and "legendre must be" in out["le"][1]


def test_addon_rejects_bad_species_and_stream():
    """legSpectralImex takes integer state indices and a BigInt stream (or
    none): fractional species and a non-BigInt stream are TypeErrors, raised
    before anything reaches the library (the handles here are never used)."""
    out = node(f"""const a = require({json.dumps(ADDON)});
      const r = [];
      const call = (species, stream) => {{
        try {{ a.legSpectralImex(1n, 2n, [3n, 4n], 5n, [0.1, 0.2], [0.3, 0.4], species, stream);
              r.push('called'); }}
        catch (e) {{ r.push([e.name, e.message]); }}
      }};
      call([0.2, 1.9], undefined);
      call([0, '1'], undefined);
      call([0, 1], 'abc');
      call([0, 1], 7);
      let se = null;
      try {{ a.legSynth(1n, [2n], 3n, {{}}); }} catch (e) {{ se = [e.name, e.message]; }}
      console.log(JSON.stringify({{r, se}}));""")
    r = out["r"]
    assert r[0][0] == r[1][0] == "TypeError" and "integer state indices" in r[0][1]
    assert r[2][0] == r[3][0] == "TypeError" and "stream must be a BigInt" in r[2][1]
    assert out["se"][0] == "TypeError" and "stream must be a BigInt" in out["se"][1]

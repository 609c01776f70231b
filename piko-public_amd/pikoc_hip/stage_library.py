"""The stage bodies the HIP kernel library implements.

pikoc compiles each stage's assignBin / schedule / process bodies to device
code (pikoc/src/pikoc.cpp:69-318, PTXBackend.cpp:356-647).  pikoc_hip binds
a pipe to the library's hand-written gfx950 kernels instead, so it may only
do so when every stage body is one those kernels implement.  This table
holds, per stage role and variant, the SHA-256 (first 16 hex digits) of the
normalised source (frontend.normalise: preprocessed with __PIKOC_DEVICE__,
comments and layout removed, macros left unexpanded) of the whole class body
and of each phase body.  The recognised bodies are the sample stages under
samples/raster and samples/reyes; tests/test_sample_stages.py runs them on
the host and checks them bit for bit against the oracle, and
tests/test_pikoc_hip.py checks that they still hash to this table.

After a deliberate change of a sample stage (and of the kernels, so that
they keep computing the same thing), refresh the table from

    python -m pikoc_hip --fingerprints -I include samples/raster/rasterPipe.h
"""
import hashlib

PHASES = ("assignBin", "schedule", "process")


def fingerprint(text):
    return None if text is None else hashlib.sha256(text.encode()).hexdigest()[:16]


def fingerprints(stage):
    """{"class": ..., "assignBin": ..., "schedule": ..., "process": ...}"""
    return {k: fingerprint(v) for k, v in stage.source.items()}


# role -> variant -> fingerprints of samples/*/ stage bodies
LIBRARY = {
    # samples/raster/vertexShader.h, VertexShaderStage
    "raster.vertex": {
        "pretransformed": {"class": "5409b41642fa7d90", "assignBin": "2e5c90b993a0f14b",
                           "schedule": "1d04f7e5f5c1ca43", "process": "017cbd695430d3b1"},
        "world": {"class": "b439ee087829f947", "assignBin": "2e5c90b993a0f14b",
                  "schedule": "1d04f7e5f5c1ca43", "process": "51b7cd2aaf8d5d0e"},
    },
    # samples/raster/raster.h, RasterStage (GORAUD)
    "raster.raster": {
        "gouraud": {"class": "4aa26a6da555a979", "assignBin": "ac826f75f3ad204e",
                    "schedule": "1d04f7e5f5c1ca43", "process": "9bc59bb970995a10"},
    },
    # samples/reyes/stages.h
    "reyes.split": {
        "split": {"class": "42c8428ab33c2f7f", "assignBin": "cc1c36a0899a9254",
                  "schedule": "1d04f7e5f5c1ca43", "process": "0ac3b61552a7e9e6"},
    },
    "reyes.dice": {
        "dice": {"class": "55d03d9eabf47032", "assignBin": "240a65374c907441",
                 "schedule": "1d04f7e5f5c1ca43", "process": "d49b3dddfe2bf420"},
    },
    "reyes.shade": {
        "shade": {"class": "d6aee88bdddef443", "assignBin": "71ddb349aff5fa8b",
                  "schedule": "1d04f7e5f5c1ca43", "process": "0d658bdf6b53b08b"},
    },
}


def match(role, stage):
    """(variant, []) when the stage's class body is a recognised one, else
    (closest variant, the phases whose bodies differ from it; ["class"] when
    only code outside the three phases differs)."""
    got = fingerprints(stage)
    best, best_diff = None, None
    for variant, want in LIBRARY[role].items():
        if got.get("class") == want["class"]:
            return variant, []
        diff = [ph for ph in PHASES if got.get(ph) != want[ph]] or ["class"]
        if best_diff is None or len(diff) < len(best_diff):
            best, best_diff = variant, diff
    return best, best_diff

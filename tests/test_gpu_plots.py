"""The measurement harness (atg.plots) in the reference's CSV schemas: every
sweep runs on a few points and writes its header and well-formed rows."""
import io

import pytest

from atg import plots

pytestmark = pytest.mark.gpu

HEADERS = {
    "build_soa": "type;duration;N",
    "build_distribution": "type;duration;distribution",
    "thread_number": "type;duration;p;N",
    "split_variants": "type;duration;splits",
    "with_transfer": "type;duration;N",
    "light_vs_heavy": "step;duration;fractionHeavy",
    "sampler_method": "type;N;numSamples;gSamplesPerSec",
    "sampling_distribution_speed": "type;N;gSamplesPerSec;distribution",
    "number_of_samples": "type;N;numSamples;gSamplesPerSec",
    "sampling_speed": "type;N;gSamplesPerSec",
    "distribution": "type;item;sampled",
    "rejection_performance": "method;N;weight0;time",
}


@pytest.mark.parametrize("sweep", sorted(HEADERS))
def test_sweep_writes_its_schema(gpu, sweep):
    out = io.StringIO()
    plots.run(sweep, out, quick=True)
    lines = out.getvalue().strip().split("\n")
    assert lines[0] == HEADERS[sweep]
    cols = lines[0].split(";")
    assert len(lines) > 1
    for row in lines[1:]:
        f = row.split(";")
        assert len(f) == len(cols), row
        for name, v in zip(cols, f):
            if name in ("duration", "gSamplesPerSec", "time"):
                assert float(v) > 0, row
    if sweep == "distribution":
        assert sum(int(r.split(";")[2]) for r in lines[1:]) == 100_000

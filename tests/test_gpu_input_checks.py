"""FASTQ validation on load: the record layout the reference's parsers assume
(genHistograms.c:60-80, metapart.c:296-369: '@' header, sequence, '+' line,
quality as long as the sequence) is checked on the device, and a record that
breaks it fails the load with MP_EINVAL naming that record."""
import numpy as np
import pytest

import metapart_amd as mp


def _records(n=64, L=120, seed=3):
    rng = np.random.default_rng(seed)
    recs = []
    for i in range(n):
        seq = "".join(rng.choice(list("ACGT"), size=L))
        recs.append(["@r%d" % i, seq, "+", "I" * L])
    return recs


def _bytes(recs):
    return ("".join("\n".join(r) + "\n" for r in recs)).encode()


def _load(data):
    buf = np.frombuffer(data, dtype=np.uint8)
    with mp.Context(k=31) as ctx:
        ctx.load(buf, [0, len(data)])
        return ctx.N


@pytest.mark.gpu
def test_well_formed_input_loads():
    print("gpu input checks: well formed", flush=True)
    assert _load(_bytes(_records())) == 64


@pytest.mark.gpu
@pytest.mark.parametrize("rec,line,text", [(0, 0, "r0"), (37, 0, "#r37"), (5, 2, "-"), (63, 2, "x"),
                                           (12, 3, "I" * 119)],
                         ids=["first_header", "header", "plus", "last_plus", "qual_length"])
def test_malformed_record_is_named(rec, line, text):
    print("gpu input checks: record %d line %d" % (rec, line), flush=True)
    recs = _records()
    recs[rec][line] = text
    with pytest.raises(mp.MPError) as e:
        _load(_bytes(recs))
    assert e.value.code == -1  # MP_EINVAL
    assert "record %d " % rec in str(e.value)

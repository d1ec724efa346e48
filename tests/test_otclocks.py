"""The per-wave timeline summary of otclocks.py on synthetic clock records
(no GPU): start and end times from s_memrealtime (100 MHz), phase durations
from s_memtime at the shader clock the records imply, waves per SIMD from the
HW_ID fields, and the lifetime of a SIMD's waves by start order."""
import otclocks


def _rec(t0, t1, clocks, cu, simd, xcc=0, se=0, chunk=0):
    r = otclocks.ClockRecord()
    r.wave = chunk
    r.realtime[0], r.realtime[1] = t0, t1
    for i, c in enumerate(clocks):
        r.clock[i] = c
    r.hw_id = (se << 13) | (cu << 8) | (simd << 4)
    r.xcc_id = xcc
    r.valid = 1
    return r


def test_summary_of_synthetic_records():
    # two SIMDs: SIMD 0 holds two waves (the first ends at 10 us, the second at
    # 20 us), SIMD 1 one wave (ends at 15 us); shader clock 2000 MHz
    recs = [
        _rec(0, 1000, [0, 4000, 18000, 19000, 20000], cu=1, simd=0),
        _rec(10, 2010, [0, 6000, 38000, 39000, 40000], cu=1, simd=0),
        _rec(0, 1500, [0, 5000, 28000, 29000, 30000], cu=2, simd=1, chunk=1),
    ]
    out = otclocks.summarise(recs)
    assert "waves 3 on 2 SIMDs (waves per SIMD: 1 x1, 2 x1)" in out
    assert "launch span 20.10 us" in out
    assert "shader clock 2000 MHz" in out
    # start order on SIMD 0: the wave that started at 0 lived 10 us, the one at 0.1 us 20 us
    assert "0: 12.50 us" in out and "1: 20.00 us" in out
    # pipeline fill of the three waves: 2, 3 and 2.5 us
    fill = next(l for l in out.splitlines() if "pipeline fill" in l)
    assert fill.split()[-1] == "3.00" and fill.split()[3] == "2.00"
    # by z-chunk: chunk 0 holds the two waves of SIMD 0 (mean 15 us, last end
    # 20.10 us) and the single last-ending wave; chunk 1 the other (15 us)
    rows = [l.split() for l in out.splitlines() if l.strip().startswith("chunk")]
    assert rows[0][1:] == ["0:", "2", "waves", "15.00", "/", "20.10", "1"]
    assert rows[1][1:] == ["1:", "1", "waves", "15.00", "/", "15.00", "0"]
    # one die (XCC 0): mean lifetime 15 us, last end 20.10 us, 2000 MHz
    assert "by XCD (XCC_ID): mean lifetime / last end, us; shader clock, MHz: 0: 15.00 / 20.10; 2000" in out

"""Call ordering on one engine (GPU): graph capture and replay keep the
spectra-ring bookkeeping right, calls on different streams are ordered by the
library, and the settings that cannot hold are refused.

* A capture runs nothing, so the eager call after it must transform the
  frames it needs again; a replay leaves the ring as its capture recorded, so
  an eager call after a replay may reuse it.  Each sequence is compared bit
  for bit with one eager call over the whole clip.
* Two ranges of one engine on two different streams, without a host wait in
  between, equal the same ranges on one stream.
"""
import numpy as np
import pytest

from harness import make_clip

pytestmark = pytest.mark.gpu

ARGS = "sigma=2;bt=3;bw=48;bh=48;ow=16;oh=16;sharpen=0.3"


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def _setup(torch, F=48, seed=12):
    from test_gpu_run import _device_clip, make_engine
    from neo_fft3d_amd import make_format
    clip = make_clip("yuv", 8, False, width=320, height=200, nframes=F, seed=seed)
    fmt = make_format("yuv", 8)
    bufs = [_device_clip(torch, clip, p) for p in range(3)]

    def engines():
        return [make_engine(ARGS, fmt, clip.width, clip.height, F, p) for p in range(3)]
    return clip, bufs, engines


def _items(engs, bufs, first, count):
    return [(e, first, count, sp, st, dp, st) for e, (src, dst, sp, dp, st) in zip(engs, bufs)]


def _reference(torch, bufs, engines, F):
    from neo_fft3d_amd import run_device_batch
    engs = engines()
    run_device_batch(_items(engs, bufs, 0, F), 0)
    engs[0].sync()
    torch.cuda.synchronize()
    want = [b[1].cpu().numpy().copy() for b in bufs]
    for b in bufs:
        b[1].zero_()
    torch.cuda.synchronize()
    return want


def _check(torch, bufs, want, label):
    torch.cuda.synchronize()
    bad = []
    for p, b in enumerate(bufs):
        got = b[1].cpu().numpy()
        d = np.abs(got.astype(int) - want[p].astype(int))
        frames = [n for n in range(d.shape[0]) if d[n].max() > 0]
        print(f"{label} plane {p}: max diff {d.max()}, frames that differ {frames}", flush=True)
        if frames:
            bad.append(p)
    assert not bad, (label, bad)


def test_capture_then_eager_next_range(torch_cuda):
    from neo_fft3d_amd import Graph, run_device_batch
    torch = torch_cuda
    clip, bufs, engines = _setup(torch)
    want = _reference(torch, bufs, engines, 48)
    engs = engines()
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    run_device_batch(_items(engs, bufs, 0, 16), s)          # pattern-free; warms the buffers
    g = Graph(_items(engs, bufs, 16, 16), s)                  # captured, not run
    run_device_batch(_items(engs, bufs, 32, 16), s)          # must not trust the capture's records
    g.launch(s)                                               # frames 16..31
    stream.synchronize()
    _check(torch, bufs, want, "capture->eager")
    g.close()


def test_eager_replay_eager(torch_cuda):
    from neo_fft3d_amd import Graph, run_device_batch
    torch = torch_cuda
    clip, bufs, engines = _setup(torch)
    want = _reference(torch, bufs, engines, 48)
    engs = engines()
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    g = Graph(_items(engs, bufs, 0, 16), s)                   # captured first
    run_device_batch(_items(engs, bufs, 16, 16), s)           # eager 16..31: ring holds 30, 31, 32
    g.launch(s)                                               # replay 0..15: ring holds 14, 15, 16
    run_device_batch(_items(engs, bufs, 32, 16), s)           # must not take 31, 32 from the ring
    stream.synchronize()
    _check(torch, bufs, want, "eager->replay->eager")
    # a replay followed by the range that continues it reuses the ring
    for b in bufs:
        b[1].zero_()
    engs2 = engines()
    g2 = Graph(_items(engs2, bufs, 0, 24), s)
    g2.launch(s)
    run_device_batch(_items(engs2, bufs, 24, 24), s)
    stream.synchronize()
    _check(torch, bufs, want, "replay->continue")
    g.close()
    g2.close()


def test_ring_bound_graph_checks_the_ring(torch_cuda):
    """A graph bound to the spectra rings (nf3d_run_device_batch_graph_ring)
    reads the frames the eager call before its capture left there: replayed
    after that call and followed by the next range, the clip equals one eager
    call bit for bit.  Once the rings hold frames that are not the capture's
    shifted alike, the replay is refused and nothing runs."""
    from neo_fft3d_amd import Graph, Nf3dError, run_device_batch
    torch = torch_cuda
    clip, bufs, engines = _setup(torch)
    want = _reference(torch, bufs, engines, 48)
    engs = engines()
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    run_device_batch(_items(engs, bufs, 0, 16), s)           # ring holds 14, 15, 16
    g = Graph(_items(engs, bufs, 16, 16), s, ring=True)       # reads 15, 16 from the ring
    g.launch(s)                                               # frames 16..31
    run_device_batch(_items(engs, bufs, 32, 16), s)           # continues from the replay's ring
    stream.synchronize()
    _check(torch, bufs, want, "ring-bound replay")
    before = [b[1].cpu().numpy().copy() for b in bufs]
    with pytest.raises(Nf3dError):                            # rings hold 45, 46, 47 now
        g.launch(s)
    stream.synchronize()
    for p, b in enumerate(bufs):
        assert np.array_equal(b[1].cpu().numpy(), before[p]), p
    g.close()


def test_out_of_order_eager_ranges(torch_cuda):
    """Ranges run out of order eagerly: every call transforms what its ring
    does not hold."""
    from neo_fft3d_amd import run_device_batch
    torch = torch_cuda
    clip, bufs, engines = _setup(torch)
    want = _reference(torch, bufs, engines, 48)
    engs = engines()
    stream = torch.cuda.Stream()
    s = stream.cuda_stream
    for first in (16, 0, 32):
        run_device_batch(_items(engs, bufs, first, 16), s)
    stream.synchronize()
    _check(torch, bufs, want, "out of order")


def test_calls_on_two_streams_are_ordered(torch_cuda):
    from neo_fft3d_amd import run_device_batch
    torch = torch_cuda
    clip, bufs, engines = _setup(torch)
    want = _reference(torch, bufs, engines, 48)
    engs = engines()
    a, b = torch.cuda.Stream(), torch.cuda.Stream()
    # busy work on stream a first so a missing wait would show
    with torch.cuda.stream(a):
        x = torch.randn(1 << 24, device="cuda")
        for _ in range(20):
            x = x * 1.0001
    run_device_batch(_items(engs, bufs, 0, 24), a.cuda_stream)
    run_device_batch(_items(engs, bufs, 24, 24), b.cuda_stream)
    b.synchronize()
    a.synchronize()
    _check(torch, bufs, want, "two streams")


def test_matrix_core_form_refuses_float(torch_cuda):
    from test_gpu_run import engine_params
    from neo_fft3d_amd import Engine, Nf3dError, make_format
    e = Engine(engine_params("sigma=3;bt=3;bw=32;bh=32"), make_format("rgb", 32, True), 96, 64, 8, 0)
    with pytest.raises(Nf3dError, match="integer formats only"):
        e.set_dft(1)
    e.close()


def test_pattern_display_capture_after_first_call(torch_cuda):
    from test_gpu_run import _device_clip, make_engine
    from neo_fft3d_amd import Graph, Nf3dError, make_format, run_device_batch
    torch = torch_cuda
    args = "sigma=3;bt=3;bw=16;bh=16;pfactor=1;pshow=1"
    clip = make_clip("gray", 8, False, width=160, height=128, nframes=8, seed=4)
    fmt = make_format("gray", 8)
    e = make_engine(args, fmt, 160, 128, 8, 0)
    src, dst, sp, dp, st = _device_clip(torch, clip, 0)
    items = [(e, 0, 8, sp, st, dp, st)]
    stream = torch.cuda.Stream()
    with pytest.raises(Nf3dError, match="pattern display"):
        Graph(items, stream.cuda_stream)
    run_device_batch(items, stream.cuda_stream)
    # the display flattens the windows for good after its first frame
    # (src/fft3d_engine.h:505-565): a replay equals a second call
    run_device_batch(items, stream.cuda_stream)
    stream.synchronize()
    want = dst.cpu().numpy().copy()
    dst.zero_()
    torch.cuda.synchronize()
    g = Graph(items, stream.cuda_stream)
    g.launch(stream.cuda_stream)
    stream.synchronize()
    assert np.array_equal(dst.cpu().numpy(), want)
    g.close()


def test_host_fed_batch_matches_device_batch(torch_cuda):
    """The host-fed pipeline (pinned staging, chunks uploaded, filtered and
    downloaded on three streams) gives the device-resident batch's output bit
    for bit, for a range starting and ending inside the clip, with crop."""
    from neo_fft3d_amd import make_format, run_device_batch, run_host_batch
    from test_gpu_run import make_engine
    torch = torch_cuda
    F = 53
    clip = make_clip("yuv", 8, False, width=320, height=200, nframes=F, seed=13)
    fmt = make_format("yuv", 8)
    for args in (ARGS + ";l=8;t=4;r=6;b=2", "sigma=3;bt=0;bw=32;bh=32;ow=10;oh=10", "sigma=3;bt=5;bw=16;bh=16"):
        first, count = 3, 45
        host_src = [np.stack([clip.frames[n][p] for n in range(F)]) for p in range(3)]
        host_dst = [np.zeros_like(h) for h in host_src]
        engs = [make_engine(args, fmt, 320, 200, F, p) for p in range(3)]
        items = [(e, first, count, [h[n].ctypes.data for n in range(F)], h.strides[1],
                  [d[n].ctypes.data for n in range(F)], d.strides[1]) for e, h, d in zip(engs, host_src, host_dst)]
        print(f"host batch {args}", flush=True)
        run_host_batch(items)
        # the device batch over the same range
        from test_gpu_run import _device_clip
        bufs = [_device_clip(torch, clip, p) for p in range(3)]
        engs2 = [make_engine(args, fmt, 320, 200, F, p) for p in range(3)]
        run_device_batch([(e, first, count, sp, st, dp, st) for e, (s_, d_, sp, dp, st) in zip(engs2, bufs)], 0)
        engs2[0].sync()
        torch.cuda.synchronize()
        for p in range(3):
            want = bufs[p][1].cpu().numpy()[first:first + count]
            got = host_dst[p][first:first + count]
            kv = dict(a.split("=") for a in args.split(";"))
            s = 1 if p else 0
            l, t, r, b = (int(kv.get(k, 0)) >> s for k in "ltrb")
            h, w = got.shape[1:]
            region = (slice(None), slice(t, h - b), slice(l, w - r))
            d = np.abs(got[region].astype(int) - want[region].astype(int))
            print(f"  plane {p}: max diff {d.max()}", flush=True)
            assert d.max() == 0, (args, p)
            if l or t or r or b:  # crop borders untouched
                mask = np.ones(got.shape, bool)
                mask[region] = False
                assert (got[mask] == 0).all()


def test_host_fed_pinned_planes_and_repeated_calls(torch_cuda):
    """Caller planes in page-locked memory are uploaded and returned straight
    (no staging copy) with the pageable path's output bit for bit, crop
    borders untouched; a second call on the same engines reuses their device
    slots and staging and repeats the output (Wiener, bt = 5)."""
    from neo_fft3d_amd import make_format, run_host_batch
    from test_gpu_run import make_engine
    torch = torch_cuda
    F = 40
    clip = make_clip("yuv", 8, False, width=320, height=200, nframes=F, seed=23)
    fmt = make_format("yuv", 8)
    args = "sigma=3;bt=5;bw=16;bh=16;l=8;t=4;r=6;b=2"
    first, count = 2, 35
    host_src = [np.stack([clip.frames[n][p] for n in range(F)]) for p in range(3)]
    page_dst = [np.zeros_like(h) for h in host_src]
    engs = [make_engine(args, fmt, 320, 200, F, p) for p in range(3)]
    items = [(e, first, count, [h[n].ctypes.data for n in range(F)], h.strides[1],
              [d[n].ctypes.data for n in range(F)], d.strides[1]) for e, h, d in zip(engs, host_src, page_dst)]
    print("host batch, pageable planes", flush=True)
    run_host_batch(items)
    again = [np.zeros_like(h) for h in host_src]
    items2 = [(e, first, count, [h[n].ctypes.data for n in range(F)], h.strides[1],
               [d[n].ctypes.data for n in range(F)], d.strides[1]) for e, h, d in zip(engs, host_src, again)]
    print("host batch, pageable planes, second call", flush=True)
    run_host_batch(items2)
    pin_src = [torch.from_numpy(h).pin_memory() for h in host_src]
    pin_dst = [torch.zeros(h.shape, dtype=torch.uint8).pin_memory() for h in host_src]
    engs3 = [make_engine(args, fmt, 320, 200, F, p) for p in range(3)]
    items3 = [(e, first, count, [s_[n].data_ptr() for n in range(F)], s_.stride(1),
               [d_[n].data_ptr() for n in range(F)], d_.stride(1)) for e, s_, d_ in zip(engs3, pin_src, pin_dst)]
    print("host batch, pinned planes", flush=True)
    run_host_batch(items3)
    for p in range(3):
        a = page_dst[p][first:first + count]
        assert np.array_equal(again[p][first:first + count], a), ("repeat", p)
        assert np.array_equal(pin_dst[p].numpy()[first:first + count], a), ("pinned", p)
        print(f"  plane {p}: pinned and repeated calls equal the pageable call", flush=True)
    for e in engs + engs3:
        e.close()

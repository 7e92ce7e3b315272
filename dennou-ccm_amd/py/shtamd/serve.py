"""Request-stream driver for the transform server (sht_server_*).

Requests follow the serving workload: truncation T21/T42/T85/T170 with
probabilities 0.3/0.3/0.3/0.1 on their quadratic Gaussian grids, 1..64 levels
(one variable, so nfield = levels), forward or inverse with equal probability,
seeded CN(0, n^-3) spectra; a forward request's grid is synthesised from its
spectra on the GPU before the stream starts.  Arrivals are Poisson (seeded
exponential gaps).  All request data is resident in HBM before the clock
starts; latency runs from the scheduled arrival to the completion the server
observes.
"""
from __future__ import annotations

import time

import numpy as np

from . import DIR_FORWARD, DIR_INVERSE, PREC_FP64, Plan, Server, ShtError, SHT_ERR_BUSY, ncoef, spec_to_real
from . import synth

GRIDS = {21: (64, 32), 42: (128, 64), 85: (256, 128), 170: (512, 256)}
MIX = ((21, 0.3), (42, 0.3), (85, 0.3), (170, 0.1))


def make_requests(n: int, seed: int = 20261202, max_levels: int = 64, mix=MIX):
    rng = np.random.default_rng(seed)
    Ts = rng.choice([t for t, _ in mix], size=n, p=[p for _, p in mix])
    Ls = rng.integers(1, max_levels + 1, size=n)
    dirs = rng.integers(0, 2, size=n)
    return [(int(T), int(L), DIR_INVERSE if d == 0 else DIR_FORWARD) for T, L, d in zip(Ts, Ls, dirs)]


class RequestData:
    """Device buffers (torch tensors) of one request: input and output."""

    def __init__(self, torch, idx, T, L, direction, synth_plans):
        im, jm = GRIDS[T]
        self.T, self.nf, self.direction = T, L, direction
        self.spec = synth.spectra(T, L, cfg=4, member=idx % 4096, var=idx // 4096)
        s_dev = torch.from_numpy(spec_to_real(self.spec)).cuda()
        if direction == DIR_INVERSE:
            self.inp = s_dev
            self.out = torch.empty((L, jm, im), dtype=torch.float64, device="cuda")
        else:
            g = torch.empty((L, jm, im), dtype=torch.float64, device="cuda")
            synth_plans[T].spec_to_grid(s_dev, g, L)
            self.inp = g
            self.out = torch.empty((ncoef(T), L, 2), dtype=torch.float64, device="cuda")

    def run_alone(self, plan):
        if self.direction == DIR_INVERSE:
            plan.spec_to_grid(self.inp, self.alone_out, self.nf)
        else:
            plan.grid_to_spec(self.inp, self.alone_out, self.nf)


def build(torch, reqs, max_levels: int = 64):
    synth_plans = {T: Plan(T, *GRIDS[T], max_levels) for T in GRIDS}
    data = [RequestData(torch, i, T, L, d, synth_plans) for i, (T, L, d) in enumerate(reqs)]
    torch.cuda.synchronize()
    for p in synth_plans.values():
        p.close()
    return data


def run_stream(torch, server: Server, data, rate: float, seed: int = 7):
    """Submits the requests at Poisson arrivals of `rate` per second; returns
    per-request latency (ms, arrival to completion), wall seconds, retries."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(data))
    arrivals = np.cumsum(gaps)
    ticket_of = {}
    lat = np.full(len(data), np.nan)
    retries = 0
    t0 = time.perf_counter()
    submit_delay = np.zeros(len(data))
    i = 0
    done = 0
    while done < len(data):
        now = time.perf_counter() - t0
        while i < len(data) and arrivals[i] <= now:
            d = data[i]
            try:
                t = server.submit(d.T, d.nf, d.direction, d.inp.data_ptr(), d.out.data_ptr())
            except ShtError as e:
                if e.code != SHT_ERR_BUSY:
                    raise
                retries += 1
                break  # not admitted: retry this request on the next pass
            ticket_of[t] = i
            submit_delay[i] = (time.perf_counter() - t0) - arrivals[i]
            i += 1
        for t, ms in server.poll():
            j = ticket_of[t]
            lat[j] = ms + 1e3 * submit_delay[j]
            done += 1
        if i < len(data):
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 2e-4:
                time.sleep(wait - 1e-4)
        else:
            time.sleep(2e-5)
    wall = time.perf_counter() - t0
    return lat, wall, retries

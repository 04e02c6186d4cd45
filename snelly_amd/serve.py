"""Served rendering: continuous batching of render jobs that arrive over time,
each rendered until its estimated error reaches the job's target.

A job is a scene at some resolution with a target relative RMS error.  It is
rendered into K independently seeded chains of its image held in one state
(snl_state_create_chains), K = 2..32 chosen so that even small images give
one launch tens of thousands of lanes; the chains' spread, reduced on the GPU
(snl_state_error), estimates the relative RMS error of their mean image.
The server loop never blocks on the GPU.  Each pass it
  1. admits waiting jobs in arrival order while the device has room for their
     states (snl_ctx_mem_info, snl_state_bytes) and the active set is not full;
  2. submits one quantum of frames for the active jobs that have no work on the
     GPU, as one snl_batch_submit, while the path samples in flight are below
     `depth` ticks' budget.  A job wants the frames its error trend asks for,
     frames_needed ~ frames_done * (error / target)^2, within (smallest,
     largest).  The tick's path-sample budget bounds each submission: jobs
     take their smallest quantum in round-robin order (the longest unserved
     first) while the budget lasts -- at least one job -- and what is left goes
     to the jobs with the least remaining work first (so a cheap job that needs
     many frames takes them in a few submissions, and a job whose error hardly
     falls does not hold back the others);
  3. polls the submitted jobs (snl_job_poll); a finished quantum updates the
     job's error estimate, and a job at or below its target retires with the
     time from arrival to completion.
So a job that arrives while others render is admitted and submitted at once,
and its launches join the others on the device (the runtime queues them on
the least loaded streams).  Pixels never read each other's state, so a job's
images do not depend on which other jobs shared the device
(tests/test_serve.py checks this).  Given the workload's program kinds
(module_scenes), the server loads their batch module and turns batch fusion
on, and the megakernel quanta of a submission run in one persistent launch.
"""
import math
import time

import numpy as np

from . import host, native


class Job:
    def __init__(self, jid, scene, target, arrival):
        self.id, self.scene, self.target, self.arrival = jid, scene, target, arrival
        self.frames = 0
        self.error = math.inf
        self.admitted = self.finished = None
        self.last_tick = -1  # the last tick that rendered the job
        self.chains = 2
        self.a = None

    @property
    def pixels(self):
        return self.scene.width * self.scene.height

    @staticmethod
    def chains_for(pixels, lanes=1 << 16, most=32):
        """the fewest chains (a power of two, 2..most) that give `lanes` lanes"""
        k = 2
        while k < most and k * pixels < lanes:
            k *= 2
        return k

    def open(self, ctx, flags, seed, lanes=1 << 16, most=32):
        s = self.scene
        self.chains = self.chains_for(self.pixels, lanes, most)
        self.a = host.Session(s, s.width, s.height, ctx, seed=seed, flags=flags, chains=self.chains)

    def close(self):
        if self.a is not None:
            self.a.close()
        self.a = None


class Server:
    """min_frames: frames per state before an error estimate counts;
    quantum: (smallest, largest) frames per state per tick;
    tick_samples: path samples per tick, smallest quanta included (None: no
    budget, every job takes what it wants each tick);
    chain_lanes, max_chains: a job gets the fewest chains (up to max_chains)
    that give it chain_lanes lanes -- a chain's frames run one after another
    in a lane, so more chains cut the time a job needs for many frames;
    reserve: device bytes kept free; max_active: cap on concurrent jobs."""

    def __init__(self, ctx, flags=native.BUILD_FAST, min_frames=4, quantum=(2, 64), reserve=8 << 30,
                 max_active=256, clock=time.perf_counter, keep_images=False, tick_samples=None,
                 chain_lanes=1 << 16, max_chains=32, depth=2, module_scenes=None):
        self.ctx, self.flags, self.keep_images = ctx, flags, keep_images
        self.min_frames, self.quantum, self.reserve, self.max_active = min_frames, quantum, reserve, max_active
        self.tick_samples = tick_samples
        self.chain_lanes, self.max_chains = chain_lanes, max_chains
        self.depth = depth
        self.module_scenes = module_scenes
        self.clock = clock
        self.ticks = 0      # submissions
        self.quanta = 0     # job quanta submitted
        self.launches = 0   # render launches they took (a persistent launch of a batch module counts once)
        self.samples = 0
        self.busy_s = 0.0   # wall time with work on the GPU

    def _need(self, job):
        """frames the job still needs by its error trend"""
        if job.frames < self.min_frames:
            return self.min_frames - job.frames
        need = job.frames * (job.error / job.target) ** 2 if job.error < math.inf else 2 * job.frames
        return max(0, math.ceil(need - job.frames))

    def _quantum(self, job):
        """the frames the job wants this tick"""
        lo, hi = self.quantum
        if job.frames < self.min_frames:
            return max(lo, self.min_frames - job.frames)
        return int(max(lo, min(hi, self._need(job))))

    def _plan(self, jobs, samples_per_frame):
        """frames per job for one tick (0: the job waits for a later tick):
        each job's wish (_quantum) when there is no budget; otherwise the
        smallest quantum (or the first min_frames) in round-robin order while
        the budget lasts, at least one job, then the rest of the budget to the
        least remaining work first"""
        want = [self._quantum(j) for j in jobs]
        if self.tick_samples is None:
            return want
        lo = self.quantum[0]
        first = [w if j.frames < self.min_frames else min(lo, w) for j, w in zip(jobs, want)]
        q = [0] * len(jobs)
        budget = self.tick_samples
        for i in sorted(range(len(jobs)), key=lambda i: (jobs[i].last_tick, jobs[i].arrival)):
            cost = first[i] * samples_per_frame[i]
            if cost > budget and any(q):
                continue
            q[i] = first[i]
            budget -= cost
        for i in sorted(range(len(jobs)), key=lambda i: max(self._need(jobs[i]), 1) * samples_per_frame[i]):
            if q[i] == 0 or budget <= 0:
                continue
            extra = min(want[i] - q[i], budget // samples_per_frame[i])
            if extra > 0:
                q[i] += extra
                budget -= extra * samples_per_frame[i]
        return q

    def run(self, arrivals, time_limit=None):
        """arrivals: [(t, scene, target)] with t in seconds from the start,
        ascending.  Returns the finished jobs (with .arrival, .admitted,
        .finished in seconds from the start) and the ones left unfinished."""
        if self.module_scenes:
            # the batch module of the workload's program kinds (from the code-object
            # cache): the megakernel quanta of each submission then run in one
            # persistent launch
            protos = [host.Session(s, 8, 8, self.ctx, flags=self.flags) for s in self.module_scenes]
            native.batch_module(self.ctx, [p.prog for p in protos])
            native.set_batch_fusion(self.ctx, True)
            for p in protos:
                p.close()
        t0 = self.clock()
        waiting = [Job(i, s, tg, t) for i, (t, s, tg) in enumerate(arrivals)]
        waiting.reverse()  # pop() takes the earliest
        queue, active, done = [], [], []
        inflight = []  # [Batch, [(job, frames)], [harvested flags]]
        busy_since = None
        while waiting or queue or active:
            now = self.clock() - t0
            if time_limit is not None and now > time_limit:
                break
            while waiting and waiting[-1].arrival <= now:
                queue.append(waiting.pop())
            # admission by free device memory
            if queue and len(active) < self.max_active:
                free, _ = native.mem_info(self.ctx)
            while queue and len(active) < self.max_active:
                job = queue[0]
                k = Job.chains_for(job.pixels, self.chain_lanes, self.max_chains)
                need = native.state_bytes(job.scene.width, job.scene.height, k)
                if host.reads_texture_unit0(job.scene):
                    need += 16 * job.pixels * k  # the per-frame radiance snapshot
                if need + self.reserve > free:
                    break
                queue.pop(0)
                job.open(self.ctx, self.flags, 1000003 * (job.id + 1), self.chain_lanes, self.max_chains)
                job.busy = False
                job.admitted = self.clock() - t0
                free -= need
                active.append(job)
            # submit the idle jobs while less than `depth` ticks of work is in flight
            launched = False
            idle = [job for job in active if not job.busy]
            pending = sum(q * job.pixels * job.chains for _, jq, got in inflight
                          for (job, q), g in zip(jq, got) if not g)
            room = self.tick_samples is None and not inflight or \
                self.tick_samples is not None and pending < self.depth * self.tick_samples
            if idle and room:
                counts = self._plan(idle, [job.pixels * job.chains for job in idle])
                jq = [(job, q) for job, q in zip(idle, counts) if q > 0]
                if jq:
                    b = native.Batch([(job.a.prog, job.a.state, q) for job, q in jq])
                    for job, _ in jq:
                        job.busy = True
                        job.last_tick = self.ticks
                    inflight.append([b, jq, [False] * len(jq)])
                    self.ticks += 1
                    launched = True
                    if busy_since is None:
                        busy_since = self.clock()
            # harvest the quanta that have finished
            harvested = False
            for entry in inflight:
                b, jq, got = entry
                for i, (job, q) in enumerate(jq):
                    if got[i] or not b.poll(i):
                        continue
                    got[i] = harvested = True
                    job.busy = False
                    job.frames += q
                    self.samples += q * job.pixels * job.chains
                    if job.frames >= self.min_frames:
                        v2, m2 = native.state_error(job.a.state)
                        job.error = math.sqrt(v2 / m2) if m2 > 0 else math.inf
                    if job.error <= job.target:
                        job.finished = self.clock() - t0
                        if self.keep_images:
                            job.image = job.a.radiance().copy()
                        active.remove(job)
                        job.close_after = True
                        done.append(job)
                if all(got):
                    self.quanta += len(jq)
                    self.launches += b.wait().trace_launches
                    b.release()
                    for job, _ in jq:
                        if getattr(job, "close_after", False):
                            job.close()
            inflight = [e for e in inflight if not all(e[2])]
            if not inflight and busy_since is not None:
                self.busy_s += self.clock() - busy_since
                busy_since = None
            if not launched and not harvested:
                # nothing to do until a quantum finishes or a job arrives
                nxt = waiting[-1].arrival - (self.clock() - t0) if waiting else 0.05
                time.sleep(max(0.0, min(nxt, 0.002 if inflight else 0.05)))
        for b, jq, got in inflight:
            b.release()
        if busy_since is not None:
            self.busy_s += self.clock() - busy_since
        for job in active:
            job.close()
        for job in done:
            job.close()
        return done, active + queue + waiting

    @staticmethod
    def report(done, left, wall_s, samples, cutoff_s=None):
        """latency percentiles over all jobs that arrived by the cutoff: a job
        left unfinished counts as right-censored at cutoff - arrival (its true
        latency is at least that).  A percentile that falls on a censored job
        is only a lower bound: `<name>_censored` says so."""
        cutoff = wall_s if cutoff_s is None else cutoff_s
        lat = [(j.finished - j.arrival, False) for j in done]
        lat += [(max(0.0, cutoff - j.arrival), True) for j in left if j.arrival <= cutoff]
        lat.sort(key=lambda x: x[0])
        out = {"jobs_done": len(done), "jobs_unfinished": sum(1 for _, c in lat if c),
               "path_samples_per_s": round(samples / wall_s, 1) if wall_s > 0 else None}
        if not lat:
            out.update({"p50_s": None, "p99_s": None, "p50_censored": False, "p99_censored": False})
            return out
        vals = np.array([v for v, _ in lat])
        for name, q in (("p50", 50), ("p99", 99)):
            k = min(len(lat) - 1, int(math.ceil(q / 100.0 * len(lat))) - 1)
            k = max(k, 0)
            out[f"{name}_s"] = round(float(vals[k]), 4)
            # exact unless a censored job (true latency >= its recorded value)
            # could rank at or below it
            out[f"{name}_censored"] = bool(any(c and v <= vals[k] for v, c in lat))
        done_lat = np.array([j.finished - j.arrival for j in done]) if done else np.array([math.nan])
        out["mean_done_s"] = round(float(done_lat.mean()), 4)
        return out

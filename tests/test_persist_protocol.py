"""The persistent batched launch's scheduling protocol (yallasql_amd/csrc/
batched.hip: ptake_job, ptake_phase, ppush and the last-finisher push in
k_persist), modelled on the CPU.  Workers are generators that yield at every
access to shared control words; a seeded scheduler interleaves them at
random, and workers leave after a bounded number of empty rounds as the
kernel's do.  Every run must execute each task exactly once, never run two
phases of one job at once, complete every job and end every worker.  The
interleavings are sequentially consistent; on the GPU the data behind each
hand-off is ordered by the release fence before every done count and the
acquire fence at the start of every task."""
import random

import pytest


def run(seed, nworkers=12, idle_limit=50):
    """One random job set and interleaving; None, or what went wrong."""
    R = random.Random(seed)
    njobs = R.randint(1, 25)
    jobs = []
    for q in range(njobs):
        split = R.random() < 0.5
        jobs.append({"split": split, "nch": R.randint(1, 5) if split else 0, "passes": R.choice([4, 8]) if split else 0,
                     "n": R.choice([0, 1, 5])})
    split_order = [q for q in range(njobs) if jobs[q]["split"] and jobs[q]["n"]]
    whole_order = [q for q in range(njobs) if not jobs[q]["split"] and jobs[q]["n"]]
    total = len(split_order) + len(whole_order)
    cap = sum(2 * jobs[q]["passes"] for q in split_order) + 1
    ctl = {"cs": 0, "cw": 0, "head": 0, "tail": 0, "done": 0, "err": 0}
    ent = [{"desc": None, "taken": 0, "done": 0} for _ in range(cap)]
    executed = []          # (job, phase, chunk)
    phase_active = {}      # job -> phase currently running (for overlap check)
    running = {}           # (job, phase) -> count in flight
    log_err = []

    def take_job(cur, order):
        if ctl[cur] >= len(order):
            return None
        yield
        i = ctl[cur]; ctl[cur] += 1
        if i >= len(order):
            return None
        return order[i]

    def take_phase():
        tail = min(ctl["tail"], cap); yield
        h = ctl["head"]; yield
        e = h
        while e < tail:
            E = ent[e]; d = E["desc"]; yield
            if d is None:
                e += 1; continue
            nt = d[2]
            if E["taken"] < nt:
                yield
                t = E["taken"]; E["taken"] += 1
                if t < nt:
                    return (e, t)
            if e == h:
                yield
                if ctl["head"] == h:
                    ctl["head"] = h + 1
                h += 1
            e += 1
        return None

    def push(job, phase, nt):
        yield
        e = ctl["tail"]; ctl["tail"] += 1
        if e >= cap:
            ctl["err"] = 1; return
        yield
        ent[e]["desc"] = (job, phase, nt)

    def worker(w, jobs_first):
        idle = 0
        while True:
            k = None
            if jobs_first:
                x = yield from take_job("cw", whole_order)
                if x is not None: k = ("job", x)
            if k is None:
                r = yield from take_phase()
                if r is not None: k = ("phase",) + r
            if k is None:
                x = yield from take_job("cs", split_order)
                if x is not None: k = ("job", x)
            if k is None:
                x = yield from take_job("cw", whole_order)
                if x is not None: k = ("job", x)
            if k is None:
                yield
                if ctl["done"] >= total:
                    return
                idle += 1
                if idle > idle_limit:
                    return
                continue
            idle = 0
            if k[0] == "job":
                j = jobs[k[1]]
                if j["split"]:
                    yield from push(k[1], 0, j["nch"])
                    continue
                for _ in range(R.randint(1, 6)): yield
                executed.append((k[1], -1, 0))
                yield
                ctl["done"] += 1
                continue
            e, t = k[1], k[2]
            job, phase, nt = ent[e]["desc"]
            # overlap check: no other phase of this job in flight
            for (jj, pp), c in running.items():
                if jj == job and pp != phase and c > 0:
                    log_err.append(("overlap", job, pp, phase))
            running[(job, phase)] = running.get((job, phase), 0) + 1
            for _ in range(R.randint(1, 6)): yield
            executed.append((job, phase, t))
            running[(job, phase)] -= 1
            yield
            d = ent[e]["done"]; ent[e]["done"] += 1
            if d == nt - 1:
                J = jobs[job]
                scatter = phase & 1; p = phase >> 1
                if not scatter or p + 1 < J["passes"]:
                    for _ in range(R.randint(0, 3)): yield
                    yield from push(job, phase + 1, J["nch"])
                else:
                    yield
                    ctl["done"] += 1

    ws = [worker(w, w % 4 == 0) for w in range(nworkers)]
    alive = list(range(nworkers))
    steps = 0
    while alive:
        i = R.choice(alive)
        try:
            next(ws[i])
        except StopIteration:
            alive.remove(i)
        steps += 1
        if steps > 10**7:
            return "steps exceeded"
    exp = set()
    for q in range(njobs):
        J = jobs[q]
        if not J["n"]: continue
        if J["split"]:
            for ph in range(2 * J["passes"]):
                for c in range(J["nch"]): exp.add((q, ph, c))
        else:
            exp.add((q, -1, 0))
    if len(executed) != len(set(executed)):
        return "duplicate task"
    if set(executed) != exp:
        return f"missing {len(exp - set(executed))} extra {len(set(executed) - exp)} (done={ctl['done']}/{total})"
    if log_err: return f"overlap {log_err[:3]}"
    if ctl["err"]: return "queue overflow"
    return None


@pytest.mark.parametrize("idle_limit", [0, 50])
def test_protocol_random_interleavings(idle_limit):
    fails = [(s, r) for s in range(300) if (r := run(s, nworkers=12, idle_limit=idle_limit))]
    assert not fails, fails[:3]


def test_protocol_at_the_gpu_worker_count():
    """768 workers: 3 workgroups on each of 256 CUs."""
    fails = [(s, r) for s in range(6) if (r := run(s, nworkers=768, idle_limit=8))]
    assert not fails, fails[:3]

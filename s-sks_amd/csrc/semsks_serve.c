/*
 * semsks_serve.c -- continuous-batching request loop over the engine.
 *
 * The reference runs one solve per process (prem_epi83.pl:81-82 starts
 * sem2d_psvsh once per event); it has no serving loop.  This loop takes a
 * stream of requests with arrival times, admits each one as a slot as soon as
 * it has arrived and its device footprint fits the engine's HBM budget
 * (first come, first served), keeps every admitted instance's wavefield
 * resident, steps all of them together (one persistent launch per time
 * step), and retires each as it reaches its own last level, which frees its
 * memory for the requests still waiting.  Each round of steps is queued on
 * the device before the round's admissions run, so admission work overlaps
 * stepping.  Latency = completion - arrival.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "semsks.h"
#include "semsks_internal.h"

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void sleep_s(double s)
{
    if (s <= 0) return;
    struct timespec ts;
    ts.tv_sec = (time_t)s;
    ts.tv_nsec = (long)((s - (double)ts.tv_sec) * 1e9);
    nanosleep(&ts, NULL);
}

static semsks_status try_admit(semsks_engine *e, semsks_request *r, int *slot)
{
    const semsks_problem *p = r->problem;
    if (r->use_planewave) {
        /* the records are generated only for a request that fits now, not
           again on every round while it waits */
        semsks_status st = semsks_room(e, p, semsks_planewave_bytes(p->nex, p->nez, p->hpn, p->nt));
        if (st != SEMSKS_OK) return st;
        st = semsks_planewave(e, p->nex, p->nez, p->hpn, p->neh, p->eh, p->nt, p->dp, r->planewave[0],
                                            r->planewave[1], r->planewave[2], r->planewave[3], NULL, NULL);
        if (st != SEMSKS_OK) return st;
    }
    return semsks_admit(e, p, -1, slot);
}

/* GLL points of one instance and the point-steps of its whole run */
static double points_of(const semsks_problem *p)
{
    return (double)(p->nex * p->hpn + 1) * (double)(p->nez * p->hpn + 1);
}
static double work_of(const semsks_problem *p) { return points_of(p) * (double)(p->nst - 1); }

semsks_status semsks_serve(semsks_engine *e, semsks_request *req, int n, int steps_per_round,
                           semsks_serve_stats *stats)
{
    semsks_serve_opts o;
    memset(&o, 0, sizeof o);
    o.steps_per_round = steps_per_round;
    o.order = SEMSKS_SERVE_FCFS;
    return semsks_serve_ex(e, req, n, &o, stats);
}

semsks_status semsks_serve_ex(semsks_engine *e, semsks_request *req, int n, const semsks_serve_opts *o,
                              semsks_serve_stats *stats)
{
    if (!e || (n > 0 && !req) || !o || o->steps_per_round < 1 ||
        (o->order != SEMSKS_SERVE_FCFS && o->order != SEMSKS_SERVE_SHORTEST) || !(o->max_active_points >= 0.0))
        return semsks_set_error(SEMSKS_EINVAL, "bad arguments");
    const int steps_per_round = o->steps_per_round;
    for (int i = 1; i < n; ++i)
        if (req[i].arrival_s < req[i - 1].arrival_s)
            return semsks_set_error(SEMSKS_EINVAL, "requests must be sorted by arrival time");
    for (int i = 0; i < n; ++i) {
        if (!req[i].problem) return semsks_set_error(SEMSKS_EINVAL, "request %d has no problem", i);
        if (req[i].use_planewave && (req[i].problem->psv || req[i].problem->sh))
            return semsks_set_error(SEMSKS_EINVAL, "request %d: plane-wave boundary and host records both given", i);
        req[i].admit_s = req[i].done_s = -1.0;
        req[i].slot = -1;
        req[i].status = SEMSKS_OK;
    }
    /* wait: requests that have arrived and are not admitted, in arrival order */
    int *wait = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *active = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!wait || !active) {
        free(wait); free(active);
        return semsks_set_error(SEMSKS_EINVAL, "out of host memory");
    }
    int nwait = 0, nact = 0, next = 0, finished = 0, max_active = 0, rounds = 0;
    double point_steps = 0.0, act_points = 0.0;
    semsks_status st = SEMSKS_OK;
    const double t0 = now_s();
    while (finished < n) {
        const double t = now_s() - t0;
        while (next < n && req[next].arrival_s <= t) wait[nwait++] = next++;
        /* the next round of the admitted instances goes out first ... */
        const int launched = nact > 0;
        if (launched) {
            st = semsks_advance(e, steps_per_round, 0);
            if (st != SEMSKS_OK) break;
            ++rounds;
        }
        /* ... and the admissions overlap it: their host work, and their
           device work on the engine's admission stream.  The next request
           is the earliest arrival (FCFS) or the one with the least work
           (SHORTEST, earliest arrival on a tie); admission stops at the
           first that does not fit, or that would take the active instances
           past max_active_points (never when none is active) */
        while (nwait > 0) {
            int w = 0;
            if (o->order == SEMSKS_SERVE_SHORTEST)
                for (int j = 1; j < nwait; ++j)
                    if (work_of(req[wait[j]].problem) < work_of(req[wait[w]].problem)) w = j;
            semsks_request *r = &req[wait[w]];
            const double pts = points_of(r->problem);
            if (o->max_active_points > 0.0 && nact > 0 && act_points + pts > o->max_active_points) break;
            int slot = -1;
            semsks_status a = try_admit(e, r, &slot);
            if ((a == SEMSKS_ENOHBM || a == SEMSKS_ENOSLOT) && nact > 0) break;
            const int k = wait[w];
            memmove(wait + w, wait + w + 1, sizeof(int) * (size_t)(nwait - w - 1));
            --nwait;
            if (a != SEMSKS_OK) {          /* cannot fit even into an idle engine, or failed */
                r->status = a;
                r->done_s = now_s() - t0;
                ++finished;
                continue;
            }
            r->slot = slot;
            r->admit_s = now_s() - t0;
            active[nact++] = k;
            act_points += pts;
            if (nact > max_active) max_active = nact;
        }
        if (!launched) {
            if (nact == 0 && next < n) sleep_s(req[next].arrival_s - (now_s() - t0));
            continue;
        }
        st = semsks_sync(e);              /* the round is done */
        if (st != SEMSKS_OK) break;
        /* harvest the instances that reached their last level */
        for (int i = 0; i < nact;) {
            semsks_request *r = &req[active[i]];
            int level = 0, done = 0;
            st = semsks_slot_status(e, r->slot, &level, &done);
            if (st != SEMSKS_OK) break;
            if (!done) { ++i; continue; }
            const semsks_problem *p = r->problem;
            if (r->traces && p->ntrace > 0) {
                st = semsks_get_traces(e, r->slot, 0, p->nst, r->traces);
                if (st == SEMSKS_ENONFINITE) { r->status = st; st = SEMSKS_OK; }   /* an unstable request */
                if (st != SEMSKS_OK) break;
            }
            r->done_s = now_s() - t0;
            point_steps += work_of(p);
            st = semsks_retire(e, r->slot);
            if (st != SEMSKS_OK) break;
            act_points -= points_of(p);
            active[i] = active[--nact];
            ++finished;
        }
        if (st != SEMSKS_OK) break;
    }
    if (stats) {
        stats->wall_s = now_s() - t0;
        stats->point_steps = point_steps;
        stats->completed = 0;
        for (int i = 0; i < n; ++i) stats->completed += req[i].status == SEMSKS_OK && req[i].done_s >= 0;
        stats->max_active = max_active;
        stats->rounds = rounds;
    }
    /* on a device error, whatever is still admitted is retired */
    for (int i = 0; i < nact; ++i) semsks_retire(e, req[active[i]].slot);
    free(wait);
    free(active);
    return st;
}

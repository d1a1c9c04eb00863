// serving.hpp — request loop (continuous batching over HIP streams with HBM-budget admission
// and resident operators) and the batched multi-problem solve.
#pragma once

#include "runtime.hpp"

struct pdehip_server;

namespace pdehip {

int solve_batch(Ctx* ctx, int count, const pdehip_problem* probs, const pdehip_options& opt,
                const pdehip_batch_config* cfg, double* const* u_out, pdehip_report* reports);

pdehip_server* server_create(Ctx* ctx, const pdehip_server_config* cfg);
int server_submit(pdehip_server* srv, const pdehip_problem& p, const pdehip_options& o, double* u_out,
                  uint64_t* ticket);
int server_poll(pdehip_server* srv, uint64_t ticket, pdehip_report* rep, bool wait);
void server_stats(pdehip_server* srv, pdehip_server_stats* st);
void server_destroy(pdehip_server* srv);

// Device bytes a solve of this problem/options needs (admission estimate).
size_t footprint_bytes(const pdehip_problem& p, const pdehip_options& o);

}  // namespace pdehip

/*
 * ops_seq_v2.h — the OPS developer-API entry header (the name apps include,
 * ops/c/include/ops_seq_v2.h) for the MI355X backend.  An app written
 * against the reference's dev API includes this file unchanged and links
 * ops_amd/libopsmi.so.  On top of include/ops_mi355x.h it provides:
 *
 *   ops_par_loop(kernel, name, block, dim, range, args...)
 *       forwarded to the hand-written ops_par_loop_<kernel> of the same
 *       kernel name — the rewrite ops_translator applies to an app
 *       (ops_translator/resources/templates/cpp/loop_host.cpp.j2:12-21), so
 *       an app may use the kernels this backend provides;
 * (The host ACC<T> the app's own kernel definitions compile against,
 * ops_execute(instance), ops_timing_output(std::ostream &) and
 * ops_checkpointing_init come from ops_mi355x.h's C++ part.)
 */
#ifndef OPS_MI355X_SEQ_V2_H
#define OPS_MI355X_SEQ_V2_H

#include <iostream>

#include "ops_mi355x.h"

#define ops_par_loop(kernel, ...) ops_par_loop_##kernel(__VA_ARGS__)

#endif

// ORACLE — test infrastructure only. Category scopes of the FLOP-counting
// build of the oracle (flop_count.h); no-ops in the ordinary build.
#pragma once
#ifdef ORACLE_FLOP_COUNT
#define ORACLE_FLOP_SCOPE(c) oracle_flop::Scope oracleFlopScope_(oracle_flop::c)
#else
#define ORACLE_FLOP_SCOPE(c) ((void)0)
#endif

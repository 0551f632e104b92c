/* Error guard of the in-process host (libtealeaf_host.so).
 *
 * The reference ends the program on every detected error: report_error
 * prints the message and calls tea_abort -> MPI_ABORT (report.f90:29-58,
 * global_mpi.f90:156-162).  The tea_leaf program keeps that behaviour.  A
 * caller that runs the host inside its own process (the Python binding,
 * benchmarks) calls the tea_host_guarded_* entry points below instead: when
 * report_error runs inside one of them, tea_host_raise returns control to
 * that entry point, which returns TEA_HOST_ERROR with the message kept for
 * tea_host_error().  The host is then in an undefined state: the caller ends
 * the run with tea_host_guarded_finalize() before it starts another.
 * Arrays that the abandoned Fortran frames had allocated are not freed. */
#include <setjmp.h>
#include <stdio.h>
#include <string.h>

#define TEA_HOST_ERROR (-2)

/* bind(C) entry points of tea_host.f90 */
int tea_host_init(const char *deck, int ndeck, const char *outp, int nout, int device);
int tea_host_step(void);
int tea_host_run(void);
int tea_host_finalize(void);

static jmp_buf *guard_env = NULL;
static char guard_msg[1024];

/* Called by report_error after it has printed the message.  Returns when no
 * guarded call is active (report_error then stops the program). */
void tea_host_raise(const char *loc, int nloc, const char *msg, int nmsg) {
  if (!guard_env) return;
  snprintf(guard_msg, sizeof guard_msg, "%.*s: %.*s", nloc < 0 ? 0 : nloc, loc,
           nmsg < 0 ? 0 : nmsg, msg);
  jmp_buf *env = guard_env;
  guard_env = NULL;
  longjmp(*env, 1);
}

const char *tea_host_error(void) { return guard_msg; }

/* setjmp must run in the frame that stays live until the guarded call
 * returns, so each entry point expands the guard itself. */
#define GUARDED(call)              \
  do {                             \
    jmp_buf env;                   \
    if (setjmp(env)) return TEA_HOST_ERROR; \
    guard_env = &env;              \
    const int rc__ = (call);       \
    guard_env = NULL;              \
    return rc__;                   \
  } while (0)

int tea_host_guarded_init(const char *deck, int ndeck, const char *outp, int nout, int device) {
  guard_msg[0] = '\0';
  GUARDED(tea_host_init(deck, ndeck, outp, nout, device));
}

int tea_host_guarded_step(void) { GUARDED(tea_host_step()); }

int tea_host_guarded_run(void) { GUARDED(tea_host_run()); }

int tea_host_guarded_finalize(void) { GUARDED(tea_host_finalize()); }

/*
 * cgfd_errors.h -- per-thread last-error message shared by the host setup and
 * the core (one library; cgfd_last_error() is defined in cgfd_errors.cpp).
 */
#ifndef CGFD_ERRORS_H
#define CGFD_ERRORS_H

/* Records a printf-style message for cgfd_last_error() and returns code. */
int cgfd_set_error(int code, const char *fmt, ...)
#if defined(__GNUC__) || defined(__clang__)
    __attribute__((format(printf, 2, 3)))
#endif
    ;

/* Clears the calling thread's message. */
void cgfd_clear_error(void);

#endif

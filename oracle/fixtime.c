/* fixtime.c -- pins time() for builds of the reference programs.
 *
 * TEST INFRASTRUCTURE ONLY (linked into oracle/_ref binaries, never into the
 * product).  The reference seeds every generator from time() (src/mc_eur.cpp:19,
 * src/mc_asia.cpp:23, include/common.h:192), so a fixed PCF_FIXTIME makes its
 * serial runs bit-reproducible.  Unset, it returns the real time. */
#include <stdlib.h>
#include <time.h>

time_t time(time_t* t) {
    const char* s = getenv("PCF_FIXTIME");
    time_t v;
    if (s && *s) {
        v = (time_t)strtoll(s, 0, 10);
    } else {
        struct timespec ts;
        clock_gettime(CLOCK_REALTIME, &ts);
        v = ts.tv_sec;
    }
    if (t) *t = v;
    return v;
}

/* timer.h -- wall clock in seconds since the first call (reference
 * timer.c:8-22; its header is not shipped).  The first call returns 0.0. */
#ifndef DM_TIMER_H
#define DM_TIMER_H
#ifdef __cplusplus
extern "C" {
#endif
double timer(void);
#ifdef __cplusplus
}
#endif
#endif

/* Counters of the host stand-in for the device library (stub_dev.cpp). */
#ifndef STUB_DEV_H
#define STUB_DEV_H
#ifdef __cplusplus
extern "C" {
#endif
int stub_live(void);      /* contexts alive                  */
int stub_created(void);   /* contexts created                */
int stub_max_batch(void); /* largest number of solves in one */
#ifdef __cplusplus
}
#endif
#endif

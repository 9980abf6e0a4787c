/* Wall clock in seconds (CLOCK_MONOTONIC_RAW), as the reference's wtime()
 * (wtime.c:6-12). */
#pragma once
double wtime(void);

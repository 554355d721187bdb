/* Prints the reference's ray-segment counter (number of localhit calls,
 * raytrace.c:22 and :606) when the process exits.  Linked only into the
 * oracle/_ref/*_count variants of the reference tools. */
#include <stdio.h>
#include <stddef.h>
extern size_t nrays;
__attribute__((destructor)) static void print_nrays(void)
{
	fprintf(stderr, "NRAYS %zu\n", nrays);
}

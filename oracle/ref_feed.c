/*
 * ref_feed.c -- input harness for the compiled reference CPU programs.
 *
 * TEST INFRASTRUCTURE ONLY.  The reference programs have no input reader:
 * each one calls srand(100) and draws its genotype matrix and phenotype
 * from rand() (GPU/v4.cpp:81-110, CPU/v2.cpp:56-84).  Linking this file
 * ahead of libc makes those rand() calls return a prepared stream, read
 * from the file named by $EPISDET_FEED (little-endian int32 values), so the
 * unmodified reference scan runs on a chosen dataset:
 *   - genotype draws: rand() % 3 == value, so a stored 0/1/2 is returned as is;
 *   - phenotype draws: a value r with (int)(N * ((double)r / RAND_MAX)) equal
 *     to the chosen case index.
 * The stream is prepared by tests/golden/make_golden.py.
 */
#include <stdio.h>
#include <stdlib.h>

static int *feed_buf;
static long feed_len;
static long feed_pos;

void srand(unsigned int seed)
{
    (void)seed;
    const char *path = getenv("EPISDET_FEED");
    if (!path) {
        fprintf(stderr, "ref_feed: EPISDET_FEED is not set\n");
        exit(3);
    }
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "ref_feed: cannot open %s\n", path);
        exit(3);
    }
    fseek(f, 0, SEEK_END);
    long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    feed_len = bytes / 4;
    feed_buf = (int *)malloc((size_t)(feed_len > 0 ? feed_len : 1) * sizeof(int));
    if (fread(feed_buf, 4, (size_t)feed_len, f) != (size_t)feed_len) {
        fprintf(stderr, "ref_feed: short read\n");
        exit(3);
    }
    fclose(f);
    feed_pos = 0;
}

int rand(void)
{
    if (feed_pos >= feed_len) {
        fprintf(stderr, "ref_feed: stream exhausted after %ld values\n", feed_pos);
        exit(3);
    }
    return feed_buf[feed_pos++];
}

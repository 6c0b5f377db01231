/*
 * Cached autotuner: the kernel configuration for a shape class is measured
 * once on the device and remembered, in memory and in a text file that later
 * processes read.
 *
 *   single products   kernel family (FMA, MFMA, split-K reduction), MFMA tile
 *                     size (large / 4x smaller) and whether K is split over
 *                     workgroups (execute.c)
 *   product groups    thin FMA tiles or 64 x 64 MFMA tiles for products whose
 *                     smaller side allows both (group.c)
 *
 * A shape class is the compute type, the precision mode, the power-of-two
 * bucket of every extent, the operand load modes and the epilogue form; its
 * key is a 64-bit FNV-1a hash of those fields.  Executors opt in with the
 * attribute TAPP_AMD_ATTR_AUTOTUNE (1: measure classes not in the cache,
 * 2: use the cache only); without it the built-in rules decide.
 *
 * Cache file: $TAPP_AMD_TUNE_CACHE, else $XDG_CACHE_HOME/tapp_amd/tune-<arch>.txt,
 * else $HOME/.cache/tapp_amd/tune-<arch>.txt.  One entry per line,
 * "<16 hex digit key> <value> <description>"; a later line for a key wins.
 * The file is appended to, never rewritten, so concurrent processes at worst
 * measure a class twice.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "internal.h"

typedef struct {
    uint64_t key;
    int value;
} tune_entry;

static tune_entry* tab;
static int ntab, cap;
static bool loaded;
static char path[1024];
static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;

uint64_t tk_tune_key(const int64_t* fields, int n)
{
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < n; i++) {
        uint64_t v = (uint64_t)fields[i];
        for (int b = 0; b < 8; b++) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    }
    return h;
}

int64_t tk_tune_bucket(int64_t x)
{
    int64_t b = 0;
    while (x > 0) { b++; x >>= 1; }
    return b;
}

static void put_locked(uint64_t key, int value)
{
    for (int i = 0; i < ntab; i++)
        if (tab[i].key == key) { tab[i].value = value; return; }
    if (ntab == cap) {
        int nc = cap ? 2 * cap : 64;
        tune_entry* t = realloc(tab, sizeof *t * (size_t)nc);
        if (!t) return;
        tab = t;
        cap = nc;
    }
    tab[ntab++] = (tune_entry){key, value};
}

static void mkdirs(const char* file)
{
    char dir[1024];
    snprintf(dir, sizeof dir, "%s", file);
    for (char* p = dir + 1; *p; p++)
        if (*p == '/') {
            *p = 0;
            if (mkdir(dir, 0755) != 0 && errno != EEXIST) return;
            *p = '/';
        }
}

static void load_locked(int device)
{
    if (loaded) return;
    loaded = true;
    const char* env = getenv("TAPP_AMD_TUNE_CACHE");
    if (env && *env) {
        snprintf(path, sizeof path, "%s", env);
    } else {
        char arch[64] = "gfx";
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            size_t i = 0;
            for (; i + 1 < sizeof arch && prop.gcnArchName[i] && prop.gcnArchName[i] != ':'; i++)
                arch[i] = prop.gcnArchName[i];
            arch[i] = 0;
        }
        (void)hipGetLastError();
        const char* xdg = getenv("XDG_CACHE_HOME");
        const char* home = getenv("HOME");
        if (xdg && *xdg) snprintf(path, sizeof path, "%s/tapp_amd/tune-%s.txt", xdg, arch);
        else if (home && *home) snprintf(path, sizeof path, "%s/.cache/tapp_amd/tune-%s.txt", home, arch);
        else path[0] = 0;
    }
    if (!path[0]) return;
    FILE* f = fopen(path, "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        unsigned long long k;
        int v;
        if (sscanf(line, "%16llx %d", &k, &v) == 2) put_locked((uint64_t)k, v);
    }
    fclose(f);
}

bool tk_tune_get(int device, uint64_t key, int* value)
{
    bool found = false;
    pthread_mutex_lock(&mu);
    load_locked(device);
    for (int i = 0; i < ntab; i++)
        if (tab[i].key == key) { *value = tab[i].value; found = true; break; }
    pthread_mutex_unlock(&mu);
    return found;
}

void tk_tune_put(int device, uint64_t key, int value, const char* desc)
{
    pthread_mutex_lock(&mu);
    load_locked(device);
    put_locked(key, value);
    if (path[0]) {
        mkdirs(path);
        FILE* f = fopen(path, "a");
        if (f) {
            fprintf(f, "%016llx %d %s\n", (unsigned long long)key, value, desc ? desc : "");
            fclose(f);
        }
    }
    pthread_mutex_unlock(&mu);
}

/* Average device time of `reps` launches after one untimed one, in
 * microseconds; negative when a launch fails. */
double tk_tune_time(struct tapp_exec* ex, int (*launch)(void* ctx), void* ctx, int reps)
{
    hipStream_t st = ex->stream;
    if (launch(ctx)) return -1.0;
    if (hipEventRecord(ex->ev0, st) != hipSuccess) return -1.0;
    for (int i = 0; i < reps; i++)
        if (launch(ctx)) return -1.0;
    if (hipEventRecord(ex->ev1, st) != hipSuccess || hipEventSynchronize(ex->ev1) != hipSuccess) return -1.0;
    float ms = 0;
    if (hipEventElapsedTime(&ms, ex->ev0, ex->ev1) != hipSuccess) return -1.0;
    return 1000.0 * ms / reps;
}

// Host build of inv_mod_plain (field.h: the scalar inversion every batched inversion
// and inverse_vt ends in) for tests/test_inverse_host.py: reads "r <hex>" / "q <hex>"
// lines (64 big-endian hex digits, 0 < x < m) and prints x^-1 mod r or mod p, one per line.
#include "field.h"

#include <cstdio>
#include <cstring>

using namespace h2;

template <class P>
static void invert(const char* hex)
{
    uint32_t x[8] = { 0 }, o[8];
    for (int i = 0; i < 8; ++i) {
        char buf[9];
        memcpy(buf, hex + 8 * (7 - i), 8);
        buf[8] = 0;
        sscanf(buf, "%x", &x[i]);
    }
    inv_mod_plain<P>(x, o);
    for (int i = 7; i >= 0; --i)
        printf("%08x", o[i]);
    printf("\n");
}

int main()
{
    char line[256];
    while (fgets(line, sizeof line, stdin)) {
        if (line[0] == 'r')
            invert<FrParams>(line + 2);
        else
            invert<FqParams>(line + 2);
    }
    return 0;
}

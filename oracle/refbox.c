/*
 * refbox.c -- test-infrastructure driver for the reference's own box-boundary routines
 * (applications/mesh/latticeBox2D/{genericBoundary,periodicX,periodicY,periodicXY}.c and
 * latticeBox3D/{genericBoundary3D,periodicXY3D,periodicZ3D,periodicXYZ3D}.c), which
 * oracle/Makefile compiles from their sources under the reference tree into
 * oracle/_ref/librefbox.so.  Those routines take the basicMesh of latticeBox2D.c /
 * latticeBox3D.c after its neighbour search (nb = the upstream node inside the box, -1
 * outside), write the boundary lists and complete the periodic links.  This driver
 * builds that basicMesh from a caller-supplied table, calls the routine and copies the
 * tables back, so tests can compare the library's and the oracle's tables with the
 * reference's bit for bit.
 *
 * The reference sources are compiled with malloc redirected to refbox_malloc, which
 * surrounds every block with guard zones: periodicX.c:33 writes bdPoints[0][-1] and one
 * element past the end, and the guards absorb and report such writes instead of letting
 * them corrupt the heap.
 *
 * Not product code: only tests load it.  The spot box (periodicXY3DSpot.c) reads its
 * spot from the case dictionary through dictIO, whose headers include <mpi.h>; it is not
 * built.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <basicMesh.h>

void genericBoundary(basicMesh* mesh, uint nx, uint ny);
void periodicX(basicMesh* mesh, uint nx, uint ny);
void periodicY(basicMesh* mesh, uint nx, uint ny);
void periodicXY(basicMesh* mesh, uint nx, uint ny);
void genericBoundary3D(basicMesh* mesh, uint nx, uint ny, uint nz);
void periodicXY3D(basicMesh* mesh, uint nx, uint ny, uint nz);
void periodicZ3D(basicMesh* mesh, uint nx, uint ny, uint nz);
void periodicXYZ3D(basicMesh* mesh, uint nx, uint ny, uint nz);

#define GUARD 4096
#define MAXBLK 64
static const unsigned char PAT = 0xA5;

static struct {
    unsigned char* base;
    size_t size;
} blk[MAXBLK];
static int nblk = 0;
static int overflow = 0;

void* refbox_malloc(size_t size) {
    if (nblk >= MAXBLK) {
        overflow = 1;
        return NULL;
    }
    unsigned char* b = (unsigned char*)malloc(size + 2 * GUARD);
    if (!b) return NULL;
    memset(b, PAT, size + 2 * GUARD);
    blk[nblk].base = b;
    blk[nblk].size = size;
    nblk++;
    return b + GUARD;
}

/* bytes outside any block's bounds that were written (guard bytes that lost the pattern) */
static long guard_writes(void) {
    long n = 0;
    for (int i = 0; i < nblk; i++) {
        const unsigned char* g0 = blk[i].base;
        const unsigned char* g1 = blk[i].base + GUARD + blk[i].size;
        for (int k = 0; k < GUARD; k++) n += (g0[k] != PAT) + (g1[k] != PAT);
    }
    return n;
}

static void release(void) {
    for (int i = 0; i < nblk; i++) free(blk[i].base);
    nblk = 0;
}

/*
 * d: 2 or 3; box: 0 generic, 1 periodicX, 2 periodicY, 3 periodicXY (2D) /
 * 0 generic, 3 periodicXY, 4 periodicZ, 5 periodicXYZ (3D), the library's box codes.
 * nb: [N][Q] node-major, in: the neighbour search's table, out: the completed table.
 * Returns the number of lists (<= maxlists) and fills sizes[b] and the lists back to
 * back in lists (capacity cap); *oob = guard bytes written outside the routine's blocks.
 * Returns -1 for an unknown box, -2 if the lists do not fit.
 */
int refbox_build(int d, int box, int nx, int ny, int nz, int Q, int32_t* nb, int maxlists, int32_t* sizes,
                 int32_t* lists, int64_t cap, long* oob) {
    const int64_t N = (int64_t)nx * ny * (d == 3 ? nz : 1);
    basicMesh mesh;
    memset(&mesh, 0, sizeof(mesh));
    mesh.nPoints = (uint)N;
    mesh.Q = (uint)Q;
    mesh.nb = (int**)malloc(sizeof(int*) * N);
    for (int64_t i = 0; i < N; i++) mesh.nb[i] = (int*)(nb + i * Q);
    overflow = 0;
    nblk = 0;
    if (d == 2) {
        switch (box) {
        case 0: genericBoundary(&mesh, nx, ny); break;
        case 1: periodicX(&mesh, nx, ny); break;
        case 2: periodicY(&mesh, nx, ny); break;
        case 3: periodicXY(&mesh, nx, ny); break;
        default: free(mesh.nb); return -1;
        }
    } else {
        switch (box) {
        case 0: genericBoundary3D(&mesh, nx, ny, nz); break;
        case 3: periodicXY3D(&mesh, nx, ny, nz); break;
        case 4: periodicZ3D(&mesh, nx, ny, nz); break;
        case 5: periodicXYZ3D(&mesh, nx, ny, nz); break;
        default: free(mesh.nb); return -1;
        }
    }
    free(mesh.nb);
    *oob = guard_writes() + (overflow ? 1 : 0);
    const int n = (int)mesh.bd.nbd;
    if (n > maxlists) {
        release();
        return -2;
    }
    int64_t at = 0;
    for (int b = 0; b < n; b++) {
        sizes[b] = (int32_t)mesh.bd.nbdelem[b];
        if (at + sizes[b] > cap) {
            release();
            return -2;
        }
        for (int i = 0; i < sizes[b]; i++) lists[at + i] = (int32_t)mesh.bd.bdPoints[b][i];
        at += sizes[b];
    }
    release();
    return n;
}

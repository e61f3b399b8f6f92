/*
 * Client side of the transfer protocol (include/cut_transfer.h): the Linux
 * form of test_app/client_lib/cuda_transfer.c:150-440.  Plain C, no GPU
 * dependency: it maps the server's shared-memory segments, copies the call's
 * parameters and input into them, and exchanges CommandPipeMessages with the
 * server over its UNIX-domain socket.  Like the reference, the calls return
 * void and report failures on stderr; cut_transfer_last_ok() tells whether
 * the last one completed.
 */
#define _GNU_SOURCE
#include <errno.h>
#include <fcntl.h>
#include <poll.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/un.h>
#include <time.h>
#include <unistd.h>

#include "transfer_common.h"

#define LIB_FN __attribute__((visibility("default")))
/* 20 minutes, as the reference client (cuda_transfer.c:10-11); the
 * CUT_TRANSFER_TIMEOUT_MS environment variable overrides it. */
#define RESPONSE_TIMEOUT_MS 1200000L

#define error_msg(...)                           \
    do {                                         \
        fprintf(stderr, "cut_transfer: ");       \
        fprintf(stderr, __VA_ARGS__);            \
        fputc('\n', stderr);                     \
    } while (0)

static SharedMemoryParams* g_params = NULL;
static void* g_data = NULL;
static size_t g_data_size = 0;
static int g_sock = -1;
static int g_last_ok = 0;

static void cleanup_shared_resources(void)
{
    if (g_data) munmap(g_data, g_data_size);
    if (g_params) munmap(g_params, sizeof(SharedMemoryParams));
    if (g_sock >= 0) close(g_sock);
    g_data = NULL;
    g_params = NULL;
    g_data_size = 0;
    g_sock = -1;
}

static int map_segment(const char* name, void** ptr, size_t* size, size_t min_size)
{
    const int fd = shm_open(name, O_RDWR, 0);
    if (fd < 0) {
        error_msg("cannot open shared memory '%s': %s (is cut_transfer_server running?)", name, strerror(errno));
        return 0;
    }
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < min_size) {
        error_msg("shared memory '%s' is smaller than %zu bytes", name, min_size);
        close(fd);
        return 0;
    }
    void* p = mmap(NULL, (size_t)st.st_size, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    close(fd);
    if (p == MAP_FAILED) {
        error_msg("cannot map shared memory '%s': %s", name, strerror(errno));
        return 0;
    }
    *ptr = p;
    *size = (size_t)st.st_size;
    return 1;
}

static int setup_shared_resources(void)
{
    if (g_sock >= 0 && g_data && g_params) return 1;
    cleanup_shared_resources();
    const char* base = getenv("CUT_TRANSFER_NAME");
    CutTransferNames names;
    if (!cut_transfer_names(base ? base : CUT_TRANSFER_DEFAULT_NAME, &names)) {
        error_msg("invalid transfer name '%s'", base ? base : "");
        return 0;
    }
    size_t psize = 0;
    if (!map_segment(names.data, &g_data, &g_data_size, 1) ||
        !map_segment(names.params, (void**)&g_params, &psize, sizeof(SharedMemoryParams))) {
        cleanup_shared_resources();
        return 0;
    }
    if (psize != sizeof(SharedMemoryParams)) {
        munmap(g_params, psize);
        g_params = NULL;
        error_msg("parameters segment has %zu bytes, expected %zu", psize, sizeof(SharedMemoryParams));
        cleanup_shared_resources();
        return 0;
    }
    g_sock = socket(AF_UNIX, SOCK_SEQPACKET | SOCK_CLOEXEC, 0);
    struct sockaddr_un addr;
    memset(&addr, 0, sizeof(addr));
    addr.sun_family = AF_UNIX;
    memcpy(addr.sun_path, names.socket_path, strlen(names.socket_path) + 1);
    if (g_sock < 0 || connect(g_sock, (struct sockaddr*)&addr, sizeof(addr)) != 0) {
        error_msg("cannot connect to '%s': %s", names.socket_path, strerror(errno));
        cleanup_shared_resources();
        return 0;
    }
    return 1;
}

static long response_timeout_ms(void)
{
    const char* s = getenv("CUT_TRANSFER_TIMEOUT_MS");
    const long v = s ? strtol(s, NULL, 10) : 0;
    return v > 0 ? v : RESPONSE_TIMEOUT_MS;
}

static long now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (long)ts.tv_sec * 1000L + ts.tv_nsec / 1000000L;
}

static int send_command(CudaCommand command, size_t data_size)
{
    CommandPipeMessage m;
    memset(&m, 0, sizeof(m));
    m.opcode = command;
    m.data_size = data_size;
    m.frame_count = 0;
    if (send(g_sock, &m, sizeof(m), MSG_NOSIGNAL) != (ssize_t)sizeof(m)) {
        error_msg("failed to send command %d: %s", (int)command, strerror(errno));
        return 0;
    }
    return 1;
}

static int wait_for_response(CommandPipeMessage* m)
{
    const long deadline = now_ms() + response_timeout_ms();
    for (;;) {
        const long left = deadline - now_ms();
        if (left <= 0) {
            error_msg("timed out waiting for the server after %ld ms", response_timeout_ms());
            return 0;
        }
        struct pollfd pfd = {g_sock, POLLIN, 0};
        const int r = poll(&pfd, 1, left > 1000 ? 1000 : (int)left);
        if (r < 0 && errno == EINTR) continue;
        if (r < 0) {
            error_msg("poll on the command socket failed: %s", strerror(errno));
            return 0;
        }
        if (r == 0) continue;
        const ssize_t n = recv(g_sock, m, sizeof(*m), 0);
        if (n == 0) {
            error_msg("the server closed the connection");
            return 0;
        }
        if (n != (ssize_t)sizeof(*m)) {
            error_msg("invalid message size %zd from the server", n);
            return 0;
        }
        return 1;
    }
}

/* One request: input already in the segments; on success copy out_bytes of
 * output from the data segment. */
static int transact(CudaCommand op, size_t in_bytes, void* out, size_t out_bytes)
{
    if (!send_command(op, in_bytes)) return 0;
    CommandPipeMessage m;
    if (!wait_for_response(&m)) return 0;
    if (m.opcode != ACK) {
        error_msg("the server refused command %d", (int)op);
        return 0;
    }
    if (!wait_for_response(&m)) return 0;
    if (m.opcode != SUCCESS) {
        error_msg("the server failed command %d", (int)op);
        return 0;
    }
    if (m.data_size != (unsigned long long)out_bytes) {
        error_msg("unexpected output size %llu from the server (expected %zu)", m.data_size, out_bytes);
        return 0;
    }
    memcpy(out, g_data, out_bytes);
    return 1;
}

static void finish(int ok)
{
    g_last_ok = ok;
    if (!ok) cleanup_shared_resources();
}

static void beamform_any(const void* data, size_t sample_bytes, const CudaBeamformerParameters* bp, float* output)
{
    g_last_ok = 0;
    const size_t data_size = (size_t)bp->rf_raw_dim[0] * bp->rf_raw_dim[1] * sample_bytes;
    const size_t out_size = (size_t)bp->output_points[0] * bp->output_points[1] * bp->output_points[2] * 2 * sizeof(float);
    if (!data || !output || data_size == 0 || out_size == 0) {
        error_msg("beamform: empty input or output");
        return;
    }
    if (!setup_shared_resources()) return;
    if (data_size > g_data_size || out_size > g_data_size) {
        error_msg("beamform: data size %zu / output size %zu exceed the shared segment (%zu)", data_size, out_size,
                  g_data_size);
        return;
    }
    memcpy(&g_params->beamformerParameters, bp, sizeof(*bp));
    memcpy(g_data, data, data_size);
    finish(transact(BEAMFORM_VOLUME, data_size, output, out_size));
}

LIB_FN void beamform_i16(const short* data, CudaBeamformerParameters bp, float* output)
{
    beamform_any(data, sizeof(short), &bp, output);
}

LIB_FN void beamform_f32(const float* data, CudaBeamformerParameters bp, float* output)
{
    beamform_any(data, sizeof(float), &bp, output);
}

LIB_FN void motion_detect_f32(const float* images, NCCMotionParameters params, float* motion_maps)
{
    g_last_ok = 0;
    const size_t data_size = (size_t)params.image_dims[0] * params.image_dims[1] * params.frame_count * sizeof(float);
    const size_t out_size =
        (size_t)params.motion_grid_dims[0] * params.motion_grid_dims[1] * 4 * sizeof(float) * params.frame_count;
    if (!images || !motion_maps || data_size == 0 || out_size == 0) {
        error_msg("motion_detect: empty input or output");
        return;
    }
    if (!setup_shared_resources()) return;
    if (data_size > g_data_size || out_size > g_data_size) {
        error_msg("motion_detect: data size %zu / output size %zu exceed the shared segment (%zu)", data_size,
                  out_size, g_data_size);
        return;
    }
    memcpy(&g_params->nccMotionParameters, &params, sizeof(params));
    memcpy(g_data, images, data_size);
    finish(transact(NCC_MOTION_DETECT, data_size, motion_maps, out_size));
}

LIB_FN void corr_images_f32(const float* template_image, const float* source_image, CorrImagesParameters params,
                            float* corr_map)
{
    g_last_ok = 0;
    if (params.template_dims[0] == 0 || params.template_dims[1] == 0 || params.source_dims[0] == 0 ||
        params.source_dims[1] == 0) {
        error_msg("corr_images: invalid dimensions");
        return;
    }
    if (params.template_dims[0] > params.source_dims[0] || params.template_dims[1] > params.source_dims[1]) {
        error_msg("corr_images: template dimensions must be <= source dimensions");
        return;
    }
    if (!template_image || !source_image || !corr_map) {
        error_msg("corr_images: null image");
        return;
    }
    const size_t tsize = (size_t)params.template_dims[0] * params.template_dims[1] * sizeof(float);
    const size_t ssize = (size_t)params.source_dims[0] * params.source_dims[1] * sizeof(float);
    const size_t out_size = ((size_t)params.source_dims[0] - params.template_dims[0] + 1) *
                            ((size_t)params.source_dims[1] - params.template_dims[1] + 1) * sizeof(float);
    if (!setup_shared_resources()) return;
    if (tsize + ssize > g_data_size || out_size > g_data_size) {
        error_msg("corr_images: data size %zu exceeds the shared segment (%zu)", tsize + ssize, g_data_size);
        return;
    }
    memcpy(&g_params->corrImagesParameters, &params, sizeof(params));
    memcpy(g_data, template_image, tsize);
    memcpy((char*)g_data + tsize, source_image, ssize);
    finish(transact(CORR_IMAGES, tsize + ssize, corr_map, out_size));
}

LIB_FN void deinit(void)
{
    cleanup_shared_resources();
}

LIB_FN int cut_transfer_last_ok(void)
{
    return g_last_ok;
}

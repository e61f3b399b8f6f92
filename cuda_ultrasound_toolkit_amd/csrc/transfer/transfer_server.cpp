// Server side of the transfer protocol (include/cut_transfer.h): the Linux
// form of test_app/src/transfer_server.cpp:150-258 and of the message loop and
// handlers of test_app/src/test_app.cpp:40-300.  It owns the data and
// parameter segments and the command socket, serves one client at a time
// (the next one after a disconnect, as the reference restarts its pipe), and
// runs each command through the library's own C++ entry points
// (cuda_toolkit::beamform / motion_detection / corr_images).  The GPU is
// first touched by the first compute command.
//
// Usage: cut_transfer_server [--name NAME] [--data-bytes N]
//                            [--idle-timeout-ms T] [--max-commands K]
// Prints one "ready" line on stdout once clients can connect; SIGINT/SIGTERM,
// the idle timeout (no client message for T ms, 0 = none) or K commands end
// it, removing the socket and both segments.
#include <fcntl.h>
#include <poll.h>
#include <signal.h>
#include <sys/mman.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/un.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <span>
#include <string>
#include <vector>

#include "cuda_toolkit.hpp"
#include "transfer_common.h"

namespace {

volatile sig_atomic_t g_stop = 0;
void on_signal(int) { g_stop = 1; }

struct Segment {
    std::string name;
    void* ptr = nullptr;
    size_t size = 0;
};

bool create_segment(Segment& s, const char* name, size_t size)
{
    s.name = name;
    shm_unlink(name);  // a stale segment of a server that did not exit cleanly
    const int fd = shm_open(name, O_CREAT | O_EXCL | O_RDWR, 0600);
    if (fd < 0) {
        std::cerr << "cut_transfer_server: cannot create shared memory " << name << ": " << std::strerror(errno) << "\n";
        return false;
    }
    if (ftruncate(fd, (off_t)size) != 0) {
        std::cerr << "cut_transfer_server: cannot size " << name << " to " << size << " bytes: " << std::strerror(errno)
                  << "\n";
        close(fd);
        shm_unlink(name);
        return false;
    }
    void* p = mmap(nullptr, size, PROT_READ | PROT_WRITE, MAP_SHARED, fd, 0);
    close(fd);
    if (p == MAP_FAILED) {
        std::cerr << "cut_transfer_server: cannot map " << name << ": " << std::strerror(errno) << "\n";
        shm_unlink(name);
        return false;
    }
    s.ptr = p;
    s.size = size;
    return true;
}

void destroy_segment(Segment& s)
{
    if (s.ptr) munmap(s.ptr, s.size);
    if (!s.name.empty()) shm_unlink(s.name.c_str());
    s.ptr = nullptr;
}

class Server {
public:
    Server(Segment& data, Segment& params) : data_(data), params_(*static_cast<SharedMemoryParams*>(params.ptr)) {}

    // Serve one connected client until it disconnects, the server is stopped,
    // or the command budget or idle timeout runs out.  Returns false to stop.
    bool serve(int fd, long idle_ms, long& commands_left)
    {
        fd_ = fd;
        auto last = std::chrono::steady_clock::now();
        while (!g_stop) {
            pollfd pfd{fd, POLLIN, 0};
            const int r = poll(&pfd, 1, 200);
            if (r < 0 && errno == EINTR) continue;
            if (r < 0) return false;
            if (r == 0) {
                if (idle_ms > 0 && std::chrono::steady_clock::now() - last > std::chrono::milliseconds(idle_ms))
                    return false;
                continue;
            }
            CommandPipeMessage m;
            const ssize_t n = recv(fd, &m, sizeof(m), 0);
            if (n == 0) return true;  // client disconnected: wait for the next one
            if (n != (ssize_t)sizeof(m)) {
                std::cerr << "cut_transfer_server: malformed message of " << n << " bytes\n";
                respond(ERR, 0);
                continue;
            }
            handle(m);
            last = std::chrono::steady_clock::now();
            if (commands_left > 0 && --commands_left == 0) return false;
        }
        return false;
    }

private:
    bool respond(CudaCommand op, unsigned long long size)
    {
        CommandPipeMessage m;
        std::memset(&m, 0, sizeof(m));
        m.opcode = op;
        m.data_size = size;
        return send(fd_, &m, sizeof(m), MSG_NOSIGNAL) == (ssize_t)sizeof(m);
    }

    // The output goes to the start of the data segment, which also holds the
    // input: results are computed into a staging buffer first, as the
    // reference's handlers do.
    void finish(bool ok, const std::vector<uint8_t>& out)
    {
        if (!ok) {
            respond(ERR, 0);
            return;
        }
        std::memcpy(data_.ptr, out.data(), out.size());
        respond(SUCCESS, out.size());
    }

    void handle(const CommandPipeMessage& m)
    {
        if (m.data_size > data_.size) {
            std::cerr << "cut_transfer_server: data size " << m.data_size << " exceeds the segment (" << data_.size
                      << ")\n";
            respond(ERR, 0);
            return;
        }
        const auto* in = static_cast<const uint8_t*>(data_.ptr);
        const size_t in_size = (size_t)m.data_size;
        // the opcode field is read as the int it is on the wire: a client may
        // send a value that is no CudaCommand, and loading that through the
        // enum type is undefined
        int opcode;
        static_assert(sizeof(opcode) == sizeof(m.opcode));
        std::memcpy(&opcode, &m.opcode, sizeof(opcode));
        switch (opcode) {
        case BEAMFORM_VOLUME: {  // test_app.cpp:101-154
            const CudaBeamformerParameters bp = params_.beamformerParameters;
            respond(ACK, 0);
            const size_t out_size = (size_t)bp.output_points[0] * bp.output_points[1] * bp.output_points[2] * 8;
            if (out_size == 0 || out_size > data_.size) {
                std::cerr << "cut_transfer_server: output of " << out_size << " bytes does not fit the segment\n";
                respond(ERR, 0);
                return;
            }
            std::vector<uint8_t> out(out_size);
            finish(cuda_toolkit::beamform(std::span<const uint8_t>(in, in_size), std::span<uint8_t>(out), bp), out);
            return;
        }
        case NCC_MOTION_DETECT: {  // test_app.cpp:156-213
            const NCCMotionParameters p = params_.nccMotionParameters;
            respond(ACK, 0);
            const size_t out_size = (size_t)p.motion_grid_dims[0] * p.motion_grid_dims[1] * 4 * sizeof(float) *
                                    p.frame_count;
            if (out_size == 0 || out_size > data_.size) {
                std::cerr << "cut_transfer_server: output of " << out_size << " bytes does not fit the segment\n";
                respond(ERR, 0);
                return;
            }
            std::vector<uint8_t> out(out_size);
            finish(cuda_toolkit::motion_detection(std::span<const uint8_t>(in, in_size), std::span<uint8_t>(out), p),
                   out);
            return;
        }
        case CORR_IMAGES: {  // test_app.cpp:215-300
            const CorrImagesParameters p = params_.corrImagesParameters;
            const uint t0 = p.template_dims[0], t1 = p.template_dims[1];
            const uint s0 = p.source_dims[0], s1 = p.source_dims[1];
            if (t0 == 0 || t1 == 0 || s0 == 0 || s1 == 0 || t0 > s0 || t1 > s1) {
                std::cerr << "cut_transfer_server: invalid corr_images dimensions\n";
                respond(ERR, 0);
                return;
            }
            const size_t tn = (size_t)t0 * t1, sn = (size_t)s0 * s1;
            if (in_size != (tn + sn) * sizeof(float)) {
                std::cerr << "cut_transfer_server: input size does not match the corr_images dimensions\n";
                respond(ERR, 0);
                return;
            }
            const size_t out_n = ((size_t)s0 - t0 + 1) * ((size_t)s1 - t1 + 1);
            if (out_n * sizeof(float) > data_.size) {
                respond(ERR, 0);
                return;
            }
            respond(ACK, 0);
            const float* tpl = reinterpret_cast<const float*>(in);
            std::vector<uint8_t> out(out_n * sizeof(float));
            const bool ok = cuda_toolkit::corr_images(std::span<const float>(tpl, tn), std::span<const float>(tpl + tn, sn),
                                                      std::span<float>(reinterpret_cast<float*>(out.data()), out_n),
                                                      p.template_dims, p.source_dims);
            finish(ok, out);
            return;
        }
        default:  // SVD_FILTER (not implemented by the reference either), unknown opcodes
            std::cerr << "cut_transfer_server: unsupported command " << opcode << "\n";
            respond(ERR, 0);
            return;
        }
    }

    Segment& data_;
    SharedMemoryParams& params_;
    int fd_ = -1;
};

}  // namespace

int main(int argc, char** argv)
{
    std::string name = CUT_TRANSFER_DEFAULT_NAME;
    size_t data_bytes = DATA_SMEM_SIZE;
    long idle_ms = 0, max_commands = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char* what) -> const char* {
            if (i + 1 >= argc) {
                std::cerr << "cut_transfer_server: " << what << " needs a value\n";
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--name") name = value("--name");
        else if (a == "--data-bytes") data_bytes = std::strtoull(value("--data-bytes"), nullptr, 10);
        else if (a == "--idle-timeout-ms") idle_ms = std::strtol(value("--idle-timeout-ms"), nullptr, 10);
        else if (a == "--max-commands") max_commands = std::strtol(value("--max-commands"), nullptr, 10);
        else {
            std::cerr << "usage: cut_transfer_server [--name NAME] [--data-bytes N] [--idle-timeout-ms T] "
                         "[--max-commands K]\n";
            return 2;
        }
    }
    CutTransferNames names;
    if (!cut_transfer_names(name.c_str(), &names) || data_bytes == 0) {
        std::cerr << "cut_transfer_server: invalid name or data size\n";
        return 2;
    }
    struct sigaction sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.sa_handler = on_signal;
    sigaction(SIGINT, &sa, nullptr);
    sigaction(SIGTERM, &sa, nullptr);

    Segment data, params;
    if (!create_segment(data, names.data, data_bytes) ||
        !create_segment(params, names.params, sizeof(SharedMemoryParams))) {
        destroy_segment(data);
        return 1;
    }
    std::memset(params.ptr, 0, sizeof(SharedMemoryParams));
    const int ls = socket(AF_UNIX, SOCK_SEQPACKET | SOCK_CLOEXEC, 0);
    sockaddr_un addr;
    std::memset(&addr, 0, sizeof(addr));
    addr.sun_family = AF_UNIX;
    std::memcpy(addr.sun_path, names.socket_path, std::strlen(names.socket_path) + 1);
    unlink(names.socket_path);
    if (ls < 0 || bind(ls, (sockaddr*)&addr, sizeof(addr)) != 0 || listen(ls, 4) != 0) {
        std::cerr << "cut_transfer_server: cannot listen on " << names.socket_path << ": " << std::strerror(errno)
                  << "\n";
        destroy_segment(data);
        destroy_segment(params);
        return 1;
    }
    std::printf("cut_transfer_server: ready name=%s data_bytes=%zu socket=%s\n", name.c_str(), data_bytes,
                names.socket_path);
    std::fflush(stdout);

    Server server(data, params);
    long commands_left = max_commands;
    auto idle_since = std::chrono::steady_clock::now();
    while (!g_stop) {
        pollfd pfd{ls, POLLIN, 0};
        const int r = poll(&pfd, 1, 200);
        if (r < 0 && errno == EINTR) continue;
        if (r < 0) break;
        if (r == 0) {
            if (idle_ms > 0 && std::chrono::steady_clock::now() - idle_since > std::chrono::milliseconds(idle_ms))
                break;
            continue;
        }
        const int fd = accept4(ls, nullptr, nullptr, SOCK_CLOEXEC);
        if (fd < 0) continue;
        const bool more = server.serve(fd, idle_ms, commands_left);
        close(fd);
        idle_since = std::chrono::steady_clock::now();
        if (!more) break;
    }
    close(ls);
    unlink(names.socket_path);
    destroy_segment(data);
    destroy_segment(params);
    return 0;
}

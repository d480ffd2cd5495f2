// A minimal WebSocket server (RFC 6455) for the interactive driver: one
// client at a time, text and binary messages, fragmented client messages,
// ping / pong and the closing handshake.  POSIX sockets only; the SHA-1 of
// the opening handshake follows FIPS 180-4 and the key encoding RFC 4648
// base64.  Server frames are unmasked and unfragmented; client frames must
// be masked (RFC 6455 section 5.1).
#pragma once

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace world_ws {

// SHA-1 digest (20 bytes) of `data`.
inline std::vector<uint8_t> sha1(const std::string& data) {
    auto rotl = [](uint32_t x, int s) { return (x << s) | (x >> (32 - s)); };
    uint32_t h[5] = {0x67452301u, 0xEFCDAB89u, 0x98BADCFEu, 0x10325476u, 0xC3D2E1F0u};
    std::vector<uint8_t> m(data.begin(), data.end());
    const uint64_t bits = static_cast<uint64_t>(data.size()) * 8u;
    m.push_back(0x80);
    while (m.size() % 64 != 56) m.push_back(0);
    for (int k = 7; k >= 0; --k) m.push_back(static_cast<uint8_t>(bits >> (8 * k)));
    for (size_t blk = 0; blk < m.size(); blk += 64) {
        uint32_t w[80];
        for (int t = 0; t < 16; ++t) {
            const uint8_t* p = &m[blk + 4 * t];
            w[t] = (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
        }
        for (int t = 16; t < 80; ++t) w[t] = rotl(w[t - 3] ^ w[t - 8] ^ w[t - 14] ^ w[t - 16], 1);
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4];
        for (int t = 0; t < 80; ++t) {
            uint32_t f, k;
            if (t < 20) {
                f = (b & c) | (~b & d);
                k = 0x5A827999u;
            } else if (t < 40) {
                f = b ^ c ^ d;
                k = 0x6ED9EBA1u;
            } else if (t < 60) {
                f = (b & c) | (b & d) | (c & d);
                k = 0x8F1BBCDCu;
            } else {
                f = b ^ c ^ d;
                k = 0xCA62C1D6u;
            }
            const uint32_t tmp = rotl(a, 5) + f + e + k + w[t];
            e = d;
            d = c;
            c = rotl(b, 30);
            b = a;
            a = tmp;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e;
    }
    std::vector<uint8_t> out(20);
    for (int i = 0; i < 5; ++i) {
        for (int k = 0; k < 4; ++k) out[4 * i + k] = static_cast<uint8_t>(h[i] >> (24 - 8 * k));
    }
    return out;
}

inline std::string base64(const std::vector<uint8_t>& in) {
    static const char* abc = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    size_t i = 0;
    for (; i + 2 < in.size(); i += 3) {
        const uint32_t v = (uint32_t(in[i]) << 16) | (uint32_t(in[i + 1]) << 8) | in[i + 2];
        out += abc[v >> 18];
        out += abc[(v >> 12) & 63];
        out += abc[(v >> 6) & 63];
        out += abc[v & 63];
    }
    if (i + 1 == in.size()) {
        const uint32_t v = uint32_t(in[i]) << 16;
        out += abc[v >> 18];
        out += abc[(v >> 12) & 63];
        out += "==";
    } else if (i + 2 == in.size()) {
        const uint32_t v = (uint32_t(in[i]) << 16) | (uint32_t(in[i + 1]) << 8);
        out += abc[v >> 18];
        out += abc[(v >> 12) & 63];
        out += abc[(v >> 6) & 63];
        out += '=';
    }
    return out;
}

// Sec-WebSocket-Accept for a client's Sec-WebSocket-Key (RFC 6455 section 4.2.2).
inline std::string accept_key(const std::string& key) {
    return base64(sha1(key + "258EAFA5-E914-47DA-95CA-C5AB0DC85B11"));
}

enum Opcode { kContinuation = 0, kText = 1, kBinary = 2, kClose = 8, kPing = 9, kPong = 10 };

class Connection {
public:
    explicit Connection(int fd) : fd_(fd) {}
    ~Connection() { close_socket(); }
    Connection(const Connection&) = delete;
    Connection& operator=(const Connection&) = delete;

    bool open() const { return fd_ >= 0; }

    // Reads the HTTP upgrade request and answers it; false (connection
    // closed) for anything that is not a WebSocket upgrade.
    bool handshake() {
        std::string req;
        char buf[1024];
        while (req.find("\r\n\r\n") == std::string::npos) {
            if (req.size() > 16384) return fail_http();
            const ssize_t r = ::recv(fd_, buf, sizeof(buf), 0);
            if (r <= 0) return fail_http();
            req.append(buf, static_cast<size_t>(r));
        }
        std::string key;
        size_t pos = req.find("\r\n");
        while (pos != std::string::npos && pos + 2 < req.size()) {
            const size_t end = req.find("\r\n", pos + 2);
            if (end == std::string::npos) break;
            const std::string line = req.substr(pos + 2, end - pos - 2);
            const size_t colon = line.find(':');
            if (colon != std::string::npos) {
                std::string name = line.substr(0, colon);
                for (char& ch : name) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
                if (name == "sec-websocket-key") {
                    size_t a = colon + 1;
                    while (a < line.size() && (line[a] == ' ' || line[a] == '\t')) ++a;
                    size_t b = line.size();
                    while (b > a && (line[b - 1] == ' ' || line[b - 1] == '\t')) --b;
                    key = line.substr(a, b - a);
                }
            }
            pos = end;
        }
        if (req.compare(0, 4, "GET ") != 0 || key.empty()) return fail_http();
        const std::string resp = "HTTP/1.1 101 Switching Protocols\r\nUpgrade: websocket\r\n"
                                 "Connection: Upgrade\r\nSec-WebSocket-Accept: " +
                                 accept_key(key) + "\r\n\r\n";
        return send_all(resp.data(), resp.size());
    }

    // Next complete text or binary message; false once the connection is
    // closed (by a close frame, an error or a protocol violation).  Pings
    // are answered on the way.
    bool read_message(std::string& out, Opcode& kind) {
        out.clear();
        bool in_message = false;
        for (;;) {
            uint8_t hdr[2];
            if (!recv_all(hdr, 2)) return closed();
            const bool fin = (hdr[0] & 0x80) != 0;
            const int op = hdr[0] & 0x0F;
            const bool masked = (hdr[1] & 0x80) != 0;
            uint64_t len = hdr[1] & 0x7F;
            if (len == 126) {
                uint8_t e[2];
                if (!recv_all(e, 2)) return closed();
                len = (uint64_t(e[0]) << 8) | e[1];
            } else if (len == 127) {
                uint8_t e[8];
                if (!recv_all(e, 8)) return closed();
                len = 0;
                for (int k = 0; k < 8; ++k) len = (len << 8) | e[k];
            }
            if (!masked || len > kMaxMessage || out.size() + len > kMaxMessage) return protocol_error();
            uint8_t mask[4];
            if (!recv_all(mask, 4)) return closed();
            std::string payload(static_cast<size_t>(len), '\0');
            if (len && !recv_all(&payload[0], static_cast<size_t>(len))) return closed();
            for (size_t i = 0; i < payload.size(); ++i) payload[i] = static_cast<char>(payload[i] ^ mask[i & 3]);
            if (op == kClose) {
                send_frame(kClose, payload.data(), payload.size() >= 2 ? 2 : 0);
                return closed();
            }
            if (op == kPing) {
                if (!fin || len > 125) return protocol_error();
                send_frame(kPong, payload.data(), payload.size());
                continue;
            }
            if (op == kPong) continue;
            if (op == kText || op == kBinary) {
                if (in_message) return protocol_error();
                kind = static_cast<Opcode>(op);
                in_message = true;
            } else if (op != kContinuation || !in_message) {
                return protocol_error();
            }
            out += payload;
            if (fin) return true;
        }
    }

    bool send_binary(const void* data, size_t n) { return send_frame(kBinary, data, n); }
    bool send_text(const std::string& s) { return send_frame(kText, s.data(), s.size()); }

private:
    static constexpr uint64_t kMaxMessage = 1u << 20;  // client messages are short commands

    bool send_frame(int op, const void* data, size_t n) {
        if (fd_ < 0) return false;
        uint8_t hdr[10];
        size_t h = 0;
        hdr[h++] = static_cast<uint8_t>(0x80 | op);
        if (n < 126) {
            hdr[h++] = static_cast<uint8_t>(n);
        } else if (n <= 0xFFFF) {
            hdr[h++] = 126;
            hdr[h++] = static_cast<uint8_t>(n >> 8);
            hdr[h++] = static_cast<uint8_t>(n);
        } else {
            hdr[h++] = 127;
            for (int k = 7; k >= 0; --k) hdr[h++] = static_cast<uint8_t>(static_cast<uint64_t>(n) >> (8 * k));
        }
        return send_all(hdr, h) && send_all(data, n);
    }
    bool send_all(const void* data, size_t n) {
        const char* p = static_cast<const char*>(data);
        while (n > 0) {
            const ssize_t w = ::send(fd_, p, n, MSG_NOSIGNAL);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) {
                close_socket();
                return false;
            }
            p += w;
            n -= static_cast<size_t>(w);
        }
        return true;
    }
    bool recv_all(void* data, size_t n) {
        char* p = static_cast<char*>(data);
        while (n > 0) {
            const ssize_t r = ::recv(fd_, p, n, 0);
            if (r < 0 && errno == EINTR) continue;
            if (r <= 0) return false;
            p += r;
            n -= static_cast<size_t>(r);
        }
        return true;
    }
    bool fail_http() {
        static const char kBad[] = "HTTP/1.1 400 Bad Request\r\nConnection: close\r\n\r\n";
        if (fd_ >= 0) (void)::send(fd_, kBad, sizeof(kBad) - 1, MSG_NOSIGNAL);
        close_socket();
        return false;
    }
    bool protocol_error() {
        static const uint8_t kCode[2] = {0x03, 0xEA};  // 1002: protocol error
        send_frame(kClose, kCode, 2);
        return closed();
    }
    bool closed() {
        close_socket();
        return false;
    }
    void close_socket() {
        if (fd_ >= 0) ::close(fd_);
        fd_ = -1;
    }

    int fd_;
};

// Listening socket on all interfaces; port 0 takes a free port (port()
// reports it).
class Listener {
public:
    explicit Listener(int port) {
        fd_ = ::socket(AF_INET, SOCK_STREAM, 0);
        if (fd_ < 0) return;
        const int one = 1;
        (void)::setsockopt(fd_, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
        sockaddr_in addr{};
        addr.sin_family = AF_INET;
        addr.sin_addr.s_addr = htonl(INADDR_ANY);
        addr.sin_port = htons(static_cast<uint16_t>(port));
        if (::bind(fd_, reinterpret_cast<sockaddr*>(&addr), sizeof(addr)) != 0 || ::listen(fd_, 4) != 0) {
            ::close(fd_);
            fd_ = -1;
            return;
        }
        socklen_t len = sizeof(addr);
        if (::getsockname(fd_, reinterpret_cast<sockaddr*>(&addr), &len) == 0) port_ = ntohs(addr.sin_port);
    }
    ~Listener() {
        if (fd_ >= 0) ::close(fd_);
    }
    Listener(const Listener&) = delete;
    Listener& operator=(const Listener&) = delete;

    bool ok() const { return fd_ >= 0; }
    int port() const { return port_; }

    // Blocks until a client connects; -1 on error.
    int accept_fd() {
        for (;;) {
            const int c = ::accept(fd_, nullptr, nullptr);
            if (c >= 0) {
                const int one = 1;
                (void)::setsockopt(c, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
                return c;
            }
            if (errno != EINTR) return -1;
        }
    }

private:
    int fd_ = -1;
    int port_ = 0;
};

}  // namespace world_ws

// File-backed segment stores (Store.h); the chunk layouts are this library's
// own: a file holds the chunks of its segments in write order, and an
// in-memory list per segment keeps (offset, length) of each chunk.
#include "Store.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace {

[[noreturn]] void IoFail(const std::string& what, const std::string& path) {
    throw std::runtime_error(what + " " + path + ": " + std::strerror(errno));
}

void MakeDirectories(const std::string& dir) {
    std::string cur;
    for (size_t i = 0; i <= dir.size(); i++) {
        if (i == dir.size() || dir[i] == '/') {
            if (!cur.empty() && cur != "." && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                IoFail("cannot create directory", cur);
        }
        if (i < dir.size()) cur += dir[i];
    }
}

// The chunks of several segments in one file.
class FileChunkStore : public StoreImpl {
public:
    FileChunkStore(int maxSegments, std::string path) : m_Path(std::move(path)), m_Segs(maxSegments) {}
    ~FileChunkStore() noexcept override { Close(true); }

    int MaxSegments() const override { return (int)m_Segs.size(); }
    bool HasData(int segment) const override { return !m_Segs[segment].chunks.empty(); }
    uint64_t Length(int segment) const override {
        uint64_t t = 0;
        for (const auto& c : m_Segs[segment].chunks) t += c.length;
        return t;
    }
    void Rewind(int segment) override { m_Segs[segment].read = 0; }

    void Write(int segment, const void* buffer, size_t size) override {
        std::lock_guard<std::mutex> lock(m_Mutex);
        if (m_Fd < 0) {  // the first chunk since the file was last emptied
            m_Fd = open(m_Path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
            if (m_Fd < 0) IoFail("cannot create", m_Path);
            m_End = 0;
        }
        const char* p = (const char*)buffer;
        size_t done = 0;
        while (done < size) {
            const ssize_t w = pwrite(m_Fd, p + done, size - done, (off_t)(m_End + done));
            if (w < 0) {
                if (errno == EINTR) continue;
                IoFail("cannot write", m_Path);
            }
            done += (size_t)w;
        }
        auto& s = m_Segs[segment];
        if (s.chunks.empty()) m_Nonempty++;
        s.chunks.push_back({m_End, size});
        m_End += size;
    }

    size_t Read(int segment, void* buffer, size_t size) override {
        auto& s = m_Segs[segment];
        if (s.read >= s.chunks.size()) return 0;
        const Chunk c = s.chunks[s.read];
        ensure(c.length <= size);
        char* p = (char*)buffer;
        size_t done = 0;
        while (done < c.length) {
            const ssize_t r = pread(m_Fd, p + done, c.length - done, (off_t)(c.offset + done));
            if (r < 0) {
                if (errno == EINTR) continue;
                IoFail("cannot read", m_Path);
            }
            if (r == 0) throw std::runtime_error("unexpected end of " + m_Path);
            done += (size_t)r;
        }
        s.read++;
        return c.length;
    }

    void Delete(int segment) override {
        std::lock_guard<std::mutex> lock(m_Mutex);
        auto& s = m_Segs[segment];
        if (!s.chunks.empty()) {
            s.chunks.clear();
            m_Nonempty--;
        }
        s.read = 0;
        if (m_Nonempty == 0) Close(true);  // every segment is gone: so is the file
    }

    void DeleteAll() override {
        std::lock_guard<std::mutex> lock(m_Mutex);
        for (auto& s : m_Segs) {
            s.chunks.clear();
            s.read = 0;
        }
        m_Nonempty = 0;
        Close(true);
    }

private:
    struct Chunk {
        uint64_t offset;
        uint64_t length;
    };
    struct Seg {
        std::vector<Chunk> chunks;
        size_t read = 0;
    };

    void Close(bool unlinkFile) {
        if (m_Fd >= 0) {
            close(m_Fd);
            m_Fd = -1;
            if (unlinkFile) unlink(m_Path.c_str());
        }
        m_End = 0;
    }

    std::string m_Path;
    std::vector<Seg> m_Segs;
    int m_Fd = -1;
    uint64_t m_End = 0;
    int m_Nonempty = 0;
    std::mutex m_Mutex;
};

// A file per segment, created on its first chunk.
class PerSegmentFileStore : public StoreImpl {
public:
    PerSegmentFileStore(int maxSegments, const std::string& directory, const std::string& prefix)
        : m_Dir(directory.empty() || directory.back() == '/' ? directory : directory + "/"),
          m_Prefix(prefix + "_"),
          m_Files(maxSegments) {
        MakeDirectories(m_Dir);
    }
    int MaxSegments() const override { return (int)m_Files.size(); }
    bool HasData(int segment) const override { return m_Files[segment] && m_Files[segment]->HasData(0); }
    uint64_t Length(int segment) const override { return m_Files[segment] ? m_Files[segment]->Length(0) : 0; }
    void Rewind(int segment) override {
        if (m_Files[segment]) m_Files[segment]->Rewind(0);
    }
    void Write(int segment, const void* buffer, size_t size) override { File(segment).Write(0, buffer, size); }
    size_t Read(int segment, void* buffer, size_t size) override {
        return m_Files[segment] ? m_Files[segment]->Read(0, buffer, size) : 0;
    }
    void Delete(int segment) override {
        std::lock_guard<std::mutex> lock(m_Mutex);
        m_Files[segment].reset();
    }

private:
    StoreImpl& File(int segment) {
        std::lock_guard<std::mutex> lock(m_Mutex);
        if (!m_Files[segment])
            m_Files[segment] = std::make_unique<FileChunkStore>(1, m_Dir + m_Prefix + std::to_string(segment));
        return *m_Files[segment];
    }
    std::string m_Dir, m_Prefix;
    std::vector<std::unique_ptr<StoreImpl>> m_Files;
    std::mutex m_Mutex;
};

// Segment s in part `pick(s)` of several stores over the same segments.
class SplitStore : public StoreImpl {
public:
    enum Kind { ROUND_ROBIN, RANGES };
    SplitStore(std::vector<std::unique_ptr<StoreImpl>> parts, Kind kind) : m_Parts(std::move(parts)), m_Kind(kind) {}
    int MaxSegments() const override { return m_Parts[0]->MaxSegments(); }
    bool HasData(int segment) const override { return Part(segment).HasData(segment); }
    uint64_t Length(int segment) const override { return Part(segment).Length(segment); }
    uint64_t TotalLength() const override {
        uint64_t t = 0;
        for (const auto& p : m_Parts) t += p->TotalLength();
        return t;
    }
    void Rewind(int segment) override { Part(segment).Rewind(segment); }
    void Write(int segment, const void* buffer, size_t size) override { Part(segment).Write(segment, buffer, size); }
    size_t Read(int segment, void* buffer, size_t size) override { return Part(segment).Read(segment, buffer, size); }
    void Delete(int segment) override { Part(segment).Delete(segment); }
    void DeleteAll() override {
        for (auto& p : m_Parts) p->DeleteAll();
    }

private:
    size_t Index(int segment) const {
        if (m_Kind == ROUND_ROBIN) return (size_t)segment % m_Parts.size();
        return (size_t)segment * m_Parts.size() / (size_t)MaxSegments();
    }
    StoreImpl& Part(int segment) { return *m_Parts[Index(segment)]; }
    const StoreImpl& Part(int segment) const { return *m_Parts[Index(segment)]; }
    std::vector<std::unique_ptr<StoreImpl>> m_Parts;
    Kind m_Kind;
};

}  // namespace

void Store::Write(int segment, const void* buffer, size_t size) {
    ensure(segment >= 0 && segment < MaxSegments());
    if (size == 0) return;
    m_Impl->Write(segment, buffer, size);
}

size_t Store::Read(int segment, void* buffer, size_t size) {
    ensure(segment >= 0 && segment < MaxSegments());
    return m_Impl->Read(segment, buffer, size);
}

Store Store::CreateSingleFileStore(int maxSegments, const std::string& filePath) {
    return Store(std::make_unique<FileChunkStore>(maxSegments, filePath));
}

Store Store::CreateMultiFileStore(int maxSegments, const std::string& directory, const std::string& prefix) {
    return Store(std::make_unique<PerSegmentFileStore>(maxSegments, directory, prefix));
}

Store Store::CreateSequentialStore(int maxSegments, std::vector<std::string> filePaths) {
    std::vector<std::unique_ptr<StoreImpl>> parts;
    for (const auto& f : filePaths) parts.push_back(std::make_unique<FileChunkStore>(maxSegments, f));
    return Store(std::make_unique<SplitStore>(std::move(parts), SplitStore::RANGES));
}

Store Store::CreateFileStore(int maxSegments, const std::string& name, StoreOptions options) {
    ensure(!options.directories.empty());
    std::vector<std::unique_ptr<StoreImpl>> perDir;
    for (const auto& dir : options.directories) {
        MakeDirectories(dir);
        if (options.filesPerPath == 0) {
            perDir.push_back(std::make_unique<PerSegmentFileStore>(maxSegments, dir, name));
        } else if (options.filesPerPath == 1) {
            perDir.push_back(std::make_unique<FileChunkStore>(maxSegments, dir + "/" + name));
        } else {
            std::vector<std::unique_ptr<StoreImpl>> files;
            for (int i = 0; i < options.filesPerPath; i++)
                files.push_back(std::make_unique<FileChunkStore>(maxSegments, dir + "/" + name + "_p" + std::to_string(i)));
            perDir.push_back(std::make_unique<SplitStore>(std::move(files), SplitStore::RANGES));
        }
    }
    return Store(std::make_unique<SplitStore>(std::move(perDir), SplitStore::ROUND_ROBIN));
}

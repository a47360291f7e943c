// xt_neighbor_cli.cpp -- drop-in xt_neighbor command line over libxtn.so.
//
// Flags, help text, input parsing (CSV, info CSV, AIRR TSV), error messages,
// exit codes (0 success, 1 error, 2 help/version) and output formats follow
// xtneighbor_streaming/src/cli.cu:10-277 and airr.cu:9-151. Input is parsed
// and validated on the host before the GPU is touched, so every parse error
// path runs without a device.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <chrono>
#include <climits>
#include <clocale>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "xtn.h"

namespace {

enum ReturnCode { SUCCESS = 0, ERROR = 1, EXIT = 2 };  // xtn.h:93

const char VERSION[] = "2.0.0\n";  // cli.cu:9
const char HELP_TEXT[] =           // cli.cu:10-27
    "xt_neighbor: perform either nearest neighbor search for CDR3 sequences or immune repertoire overlap using GPU-based xt_neighbor algorithm.\n"
    "\t====================\n\t Common Options\n\t====================\n"
    "\t -d or --distance [number]: distance threshold defining the neighbor (default to 1)\n"
    "\t -o or --output-path [str]: path of the output file (default to no output)\n"
    "\t -m or --measurement [leven|hamming]: distance measurement (default to leven)\n"
    "\t -v or --version: print the version of the program then exit\n"
    "\t -h or --help: print the help text of the program then exit\n"
    "\t -V or --verbose: print extra detail as the program runs for debugging purpose\n"
    "\t -a or --airr: use AIRR format for input-path instead. Relevant fields are cdr3_aa and duplicate_count\n"
    "\t====================\n\t Nearest Neighbor Options\n\t====================\n"
    "\t -i or --input-path [str] (required): path of csv input file containing exactly 1 column: CDR3 amino acid sequences\n"
    "\t -n or --input-length [number] (required): number of rows given in the input file\n"
    "\t====================\n\t Repertoire Overlap Options\n\t====================\n"
    "\t -i or --input-path [str] (required): path of csv input file containing exactly 2 columns: CDR3 amino acid sequences and their frequency. Note that the sequences are assumed to be unique\n"
    "\t -n or --input-length [number] (required): number of sequences given in the input file\n"
    "\t -I or --info-path [str] (required): path of csv input file containing exactly 1 column: repertoire sizes. Note that the order of input sequence must be sorted according to this repertoire info\n"
    "\t -N or --info-length [number] (required): number of repertoires given in the info file\n";

struct Args {  // XTNArgs, xtn.h:50-60
	int distance = 1;
	int verbose = 0;
	char* seqPath = nullptr;
	int seqLen = 0;
	char* outputPath = nullptr;
	int measure = XTN_LEVENSHTEIN;
	char* infoPath = nullptr;
	int infoLen = 0;
	int airr = 0;
};

int print_err(const char* s) {  // util.cu:13-16
	fprintf(stderr, "Error: %s\n", s);
	return ERROR;
}

int print_err_line(const char* s, int line) {  // util.cu:18-22
	fprintf(stderr, "Error: %s\n", s);
	fprintf(stderr, "Error line: %d\n", line);
	return ERROR;
}

// codec.cu:25-38
int char_encode(char c) {
	if (c < 'A' || c > 'Y')
		return -1;
	if (c == 'B' || c == 'J' || c == 'O' || c == 'U' || c == 'X')
		return -1;
	return c - '@';
}

// SeqArray (xtn.h:95-114, codec.cu:168-196): concatenated residues + offsets
struct SeqStore {
	std::vector<char> chars;
	std::vector<uint32_t> offsets{0};
	int maxSize = 0;
	int size() const { return (int)offsets.size() - 1; }

	// codec.cu:175-196: a field ends at NUL, LF, ',' or TAB, at most 500
	// residues; an empty field is accepted and stores nothing
	int append(const char* s) {
		if (size() >= maxSize)
			return print_err("SeqArray exceeding max length");
		int i = 0;
		while (i < 500) {
			char c = s[i];
			if (c == '\0' || c == '\n' || c == ',' || c == '\t')
				break;
			i++;
			if (char_encode(c) == -1)
				return print_err("invalid char");
			chars.push_back(c);
		}
		if (i == 0)
			return SUCCESS;
		offsets.push_back(offsets.back() + (uint32_t)i);
		return SUCCESS;
	}
};

// cli.cu:29-89
int parse_args(int argc, char** argv, Args* a) {
	for (int i = 1; i < argc; i++) {
		const char* cur = argv[i];
		auto value = [&](void) -> char* { return i + 1 < argc ? argv[++i] : nullptr; };
		if (!strcmp(cur, "-v") || !strcmp(cur, "--version")) {
			printf("%s", VERSION);
			return EXIT;
		} else if (!strcmp(cur, "-h") || !strcmp(cur, "--help")) {
			printf("%s", HELP_TEXT);
			return EXIT;
		} else if (!strcmp(cur, "-V") || !strcmp(cur, "--verbose")) {
			a->verbose = 1;
		} else if (!strcmp(cur, "-a") || !strcmp(cur, "--airr")) {
			a->airr = 1;
		} else if (!strcmp(cur, "-i") || !strcmp(cur, "--input-path")) {
			a->seqPath = value();
		} else if (!strcmp(cur, "-I") || !strcmp(cur, "--info-path")) {
			a->infoPath = value();
		} else if (!strcmp(cur, "-o") || !strcmp(cur, "--output-path")) {
			a->outputPath = value();
		} else if (!strcmp(cur, "-d") || !strcmp(cur, "--distance")) {
			char* v = value();
			if (!v)
				return print_err("distance must be a valid number larger than 0");
			a->distance = atoi(v);
			if (a->distance < 1)
				return print_err("distance must be a valid number larger than 0");
		} else if (!strcmp(cur, "-n") || !strcmp(cur, "--input-length")) {
			char* v = value();
			a->seqLen = v ? atoi(v) : 0;
			if (a->seqLen == 0)
				return print_err("invalid input length");
		} else if (!strcmp(cur, "-N") || !strcmp(cur, "--info-length")) {
			char* v = value();
			a->infoLen = v ? atoi(v) : 0;
			if (a->infoLen == 0)
				return print_err("invalid info length");
		} else if (!strcmp(cur, "-m") || !strcmp(cur, "--measurement")) {
			char* m = value();
			if (m && !strcmp(m, "leven"))
				a->measure = XTN_LEVENSHTEIN;
			else if (m && !strcmp(m, "hamming"))
				a->measure = XTN_HAMMING;
			else
				return print_err("invalid measure option");
		} else {
			return print_err("unknown option");
		}
	}
	if (a->seqPath == nullptr)
		return print_err("missing path for seq");
	if (a->seqLen == 0)
		return print_err("missing length for seq");
	if ((a->infoPath == nullptr) != (a->infoLen == 0))
		return print_err("repertiore path or repertoire count is missing in overlap mode");
	return SUCCESS;
}

// cli.cu:101-147
int parse_input(const char* path, SeqStore* seqs, std::vector<xtn_seq_info>* info, int len,
                bool doubleCol) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("input file reading failed");
	const int BUFFER_SIZE = 500;
	char line[BUFFER_SIZE];
	int lineNumber = 1, inputCount = 0;
	if (!fgets(line, BUFFER_SIZE, f))
		line[0] = '\0';  // header
	while (fgets(line, BUFFER_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		if (seqs->append(line) == ERROR) {
			fclose(f);
			return print_err_line("input parsing error", lineNumber);
		}
		if (doubleCol) {
			char* line2 = strchr(line, ',');
			if (!line2) {
				fclose(f);
				return print_err_line("input parsing error (comma expected)", lineNumber);
			}
			long temp = strtol(line2 + 1, nullptr, 10);
			if (temp == 0 || temp > INT_MAX || temp < INT_MIN) {
				fclose(f);
				return print_err_line("input parsing error (invalid number)", lineNumber);
			}
			if (inputCount < (int)info->size())
				(*info)[inputCount].frequency = (int32_t)temp;
		}
		inputCount++;
	}
	fclose(f);
	// a line whose first field is empty is counted without storing a
	// sequence; the reference then runs on uninitialised rows, here it is
	// reported as a length mismatch
	if (inputCount != len || seqs->size() != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

// parse_input on host threads, for files where it is exactly the serial
// parse: every line shorter than the 499 characters one fgets call takes (a
// longer line would be split into several), no NUL byte, no error, no more
// sequences than announced, and (two columns) a frequency field that strtol
// reads within its line. Returns SUCCESS or ERROR as parse_input would, or
// FALLBACK when the serial parser must decide (it then prints the reference's
// exact message and line number). The file is mapped; each thread takes a
// range of whole lines, counts and checks them, then writes its residues and
// offsets straight into the store at its prefix.
constexpr int FALLBACK = -1;

struct Residues {
	bool ok[256];
	Residues() {
		for (int c = 0; c < 256; ++c)
			ok[c] = char_encode((char)c) != -1;
	}
};

// strtol(p, nullptr, 10) for a field that ends at `end` (a line's newline):
// false when strtol could read past `end` (a field of white space or a bare
// sign, after which strtol would skip the newline into the next line)
bool field_long(const char* p, const char* end, long* out) {
	while (p < end && (*p == ' ' || *p == '\t' || *p == '\v' || *p == '\f' || *p == '\r'))
		++p;
	bool neg = false;
	if (p < end && (*p == '+' || *p == '-'))
		neg = *p++ == '-';
	if (p >= end || *p < '0' || *p > '9')
		return false;
	long v = 0;
	bool over = false;
	for (; p < end && *p >= '0' && *p <= '9'; ++p) {
		const int d = *p - '0';
		if (v > (LONG_MAX - d) / 10)
			over = true;
		else
			v = v * 10 + d;
	}
	*out = over ? (neg ? LONG_MIN : LONG_MAX) : (neg ? -v : v);
	return true;
}

int parse_input_parallel(const char* path, SeqStore* seqs, std::vector<xtn_seq_info>* info, int len,
                         bool doubleCol) {
	static const Residues R;
	const int fd = open(path, O_RDONLY);
	if (fd < 0)
		return FALLBACK;
	struct stat sb;
	if (fstat(fd, &sb) != 0 || sb.st_size <= 0) {
		close(fd);
		return FALLBACK;
	}
	const size_t size = (size_t)sb.st_size;
	void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
	close(fd);
	if (m == MAP_FAILED)
		return FALLBACK;
	struct Unmap {
		void* m;
		size_t n;
		~Unmap() { munmap(m, n); }
	} unmap{m, size};
	const char* b = static_cast<const char*>(m);
	if (memchr(b, '\0', size))
		return FALLBACK;
	// the header: the first fgets line
	const char* nl = (const char*)memchr(b, '\n', size);
	const size_t body = nl ? (size_t)(nl - b) + 1 : size;
	if ((nl ? (size_t)(nl - b) : size) >= 499)
		return FALLBACK;
	int T = (int)std::thread::hardware_concurrency();
	T = T < 1 ? 1 : (T > 16 ? 16 : T);
	if (size - body < ((size_t)1 << 20))
		T = 1;
	std::vector<size_t> cut(T + 1, size);  // chunk starts at line starts
	cut[0] = body;
	for (int t = 1; t < T; ++t) {
		size_t p = body + (size - body) * (size_t)t / (size_t)T;
		if (p < cut[t - 1])
			p = cut[t - 1];
		const char* q = p < size ? (const char*)memchr(b + p, '\n', size - p) : nullptr;
		cut[t] = q ? (size_t)(q - b) + 1 : size;
	}
	struct Part {
		size_t seqs = 0, chars = 0;
		int counted = 0;
		bool bad = false;
	};
	std::vector<Part> parts(T);
	// walks the lines of chunk t; with `out`, writes them from the part's bases
	auto scan = [&](int t, bool write, size_t seq_base, size_t char_base, int count_base) {
		Part& pt = parts[t];
		size_t a = cut[t];
		const size_t end = cut[t + 1];
		size_t ns = 0, nc = 0;
		int cnt = 0;
		while (a < end) {
			const char* e = (const char*)memchr(b + a, '\n', end - a);
			const size_t stop = e ? (size_t)(e - b) : end;  // the line is [a, stop)
			const size_t ll = stop - a;
			const bool has_nl = e != nullptr;
			const size_t next = has_nl ? stop + 1 : end;
			if (!write && ll >= 499) {
				pt.bad = true;
				return;
			}
			if (!has_nl && ll == 0)
				break;  // nothing after the last newline
			if (has_nl && (ll == 0 || (ll == 1 && b[a] == ' '))) {  // "\n" or " \n"
				a = next;
				continue;
			}
			size_t i = a;
			while (i < stop && b[i] != ',' && b[i] != '\t') {
				if (!write && !R.ok[(unsigned char)b[i]]) {
					pt.bad = true;
					return;
				}
				++i;
			}
			const size_t fl = i - a;
			if (fl > 0) {
				if (write) {
					memcpy(seqs->chars.data() + char_base + nc, b + a, fl);
					seqs->offsets[seq_base + ns + 1] = (uint32_t)(char_base + nc + fl);
				}
				++ns;
				nc += fl;
			}
			if (doubleCol) {
				const char* comma = (const char*)memchr(b + a, ',', ll);
				long temp = 0;
				if (!comma || !field_long(comma + 1, b + stop, &temp) || temp == 0 || temp > INT_MAX ||
				    temp < INT_MIN) {
					pt.bad = true;
					return;
				}
				const int row = count_base + cnt;
				if (write && row < (int)info->size())
					(*info)[row].frequency = (int32_t)temp;
			}
			++cnt;
			a = next;
		}
		if (!write) {
			pt.seqs = ns;
			pt.chars = nc;
			pt.counted = cnt;
		}
	};
	auto run = [&](auto fn) {
		if (T == 1) {
			fn(0);
			return;
		}
		std::vector<std::thread> th;
		for (int t = 0; t < T; ++t)
			th.emplace_back(fn, t);
		for (auto& x : th)
			x.join();
	};
	run([&](int t) { scan(t, false, 0, 0, 0); });
	std::vector<size_t> sbase(T + 1, 0), cbase(T + 1, 0);
	std::vector<int> nbase(T + 1, 0);
	for (int t = 0; t < T; ++t) {
		if (parts[t].bad)
			return FALLBACK;
		sbase[t + 1] = sbase[t] + parts[t].seqs;
		cbase[t + 1] = cbase[t] + parts[t].chars;
		nbase[t + 1] = nbase[t] + parts[t].counted;
	}
	const size_t stored = sbase[T];
	if (stored > (size_t)seqs->maxSize || cbase[T] > (size_t)UINT32_MAX)
		return FALLBACK;
	seqs->chars.resize(cbase[T]);
	seqs->offsets.resize(stored + 1);
	seqs->offsets[0] = 0;
	run([&](int t) { scan(t, true, sbase[t], cbase[t], nbase[t]); });
	if (nbase[T] != len || (int)stored != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

// cli.cu:152-193
int parse_info(const char* path, std::vector<xtn_seq_info>* result, int len, int seqLen) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("info file reading failed");
	const int BUFFER_SIZE = 500;
	char line[BUFFER_SIZE];
	int lineNumber = 1, inputCount = 0, resultIndex = 0;
	if (!fgets(line, BUFFER_SIZE, f))
		line[0] = '\0';
	while (fgets(line, BUFFER_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		long temp = strtol(line, nullptr, 10);
		if (temp == 0 || temp > INT_MAX || temp < INT_MIN) {
			fclose(f);
			return print_err_line("info parsing error (invalid number)", lineNumber);
		}
		if (resultIndex + temp > seqLen) {
			fclose(f);
			return print_err("total repertoires' size does not match sequence count [1]");
		}
		for (long i = 0; i < temp; i++) {
			(*result)[resultIndex].repertoire = inputCount;
			(*result)[resultIndex].original_index = resultIndex;
			resultIndex++;
		}
		inputCount++;
	}
	fclose(f);
	if (inputCount != len)
		return print_err("info length doesn't match with the actual");
	if (resultIndex != seqLen)
		return print_err("total repertoires' size does not match sequence count [2]");
	return SUCCESS;
}

// airr.cu:24-36: column start offsets of a TAB separated line
int split_offset(const char* str, int* out, int nColumn) {
	char c;
	int i = 0, count = 1;
	out[0] = 0;
	while ((c = str[i++]) != '\0') {
		if (c == '\t') {
			out[count++] = i;
			if (count == nColumn)
				return count;
		}
	}
	return count;
}

// airr.cu:41-48: the column whose start is the first occurrence of needle
int find_idx(const char* haystack, const char* needle, const int* offsets, int n) {
	const char* p = strstr(haystack, needle);
	if (!p)
		return -1;
	int pos = (int)(p - haystack);
	for (int i = 0; i < n; i++)
		if (offsets[i] == pos)
			return i;
	return -1;
}

const int DEFAULT_MAX_COLUMN = 300;  // airr.cu:11
const int DEFAULT_LINE_SIZE = 5000;  // airr.cu:13

// airr.cu:53-66
int find_header(const char* line, int& cdrIndex, int& freqIndex, bool doubleCol) {
	std::vector<int> offsets(DEFAULT_MAX_COLUMN);
	int n = split_offset(line, offsets.data(), DEFAULT_MAX_COLUMN);
	cdrIndex = find_idx(line, "cdr3_aa", offsets.data(), n);
	freqIndex = doubleCol ? find_idx(line, "duplicate_count", offsets.data(), n) : -1;
	return n;
}

// airr.cu:71-105. A header without the needed column (index -1) makes every
// row an error here; the reference indexes offsets[-1] in that case.
int extract_data(const char* line, int cdrIndex, int freqIndex, int nColumn, int lineNumber,
                 SeqStore* seqs, int& freqOut, bool doubleCol) {
	std::vector<int> offsets(nColumn > 0 ? nColumn : 1);
	int n = split_offset(line, offsets.data(), nColumn);
	if (cdrIndex < 0 || n <= cdrIndex)
		return ERROR;
	if (seqs->append(line + offsets[cdrIndex]) == ERROR)
		return print_err_line("invalid seq", lineNumber);
	if (doubleCol) {
		if (freqIndex < 0 || n <= freqIndex)
			return ERROR;
		long temp = strtol(line + offsets[freqIndex], nullptr, 10);
		if (temp == 0 || temp > INT_MAX || temp < INT_MIN)
			return print_err_line("invalid number", lineNumber);
		freqOut = (int)temp;
	} else {
		freqOut = -1;
	}
	return SUCCESS;
}

// airr.cu:116-151
int parse_airr_input(const char* path, SeqStore* seqs, std::vector<xtn_seq_info>* info, int len,
                     bool doubleCol) {
	FILE* f = fopen(path, "r");
	if (!f)
		return print_err("input file reading failed");
	std::vector<char> buf(DEFAULT_LINE_SIZE);
	char* line = buf.data();
	int lineNumber = -1, inputCount = 0, seqIndex = -1, freqIndex = -1;
	if (!fgets(line, DEFAULT_LINE_SIZE, f))
		line[0] = '\0';
	int nColumn = find_header(line, seqIndex, freqIndex, doubleCol);
	while (fgets(line, DEFAULT_LINE_SIZE, f)) {
		lineNumber++;
		if (!strcmp(line, "\n") || !strcmp(line, " \n"))
			continue;
		int freq = 0;
		if (extract_data(line, seqIndex, freqIndex, nColumn, lineNumber, seqs, freq, doubleCol) != SUCCESS) {
			fclose(f);
			return print_err_line("line parsing error", lineNumber);
		}
		if (doubleCol && inputCount < (int)info->size())
			(*info)[inputCount].frequency = freq;
		inputCount++;
	}
	fclose(f);
	if (inputCount != len || seqs->size() != len)
		return print_err("input length doesn't match with the actual");
	return SUCCESS;
}

// util.cu:40-50
void print_args(const Args& a) {
	printf("XTNArgs{\n");
	printf("\tdistance: %d\n", a.distance);
	printf("\tverbose: %d\n", a.verbose);
	printf("\tseqLen: %'d\n", a.seqLen);
	printf("\tseqPath: \"%s\"\n", a.seqPath ? a.seqPath : "(null)");
	printf("\toutputPath: \"%s\"\n", a.outputPath ? a.outputPath : "(null)");
	printf("\tmeasure: \"%s\"\n", (a.measure == XTN_LEVENSHTEIN) ? "leven" : "hamming");
	printf("\tinfoPath: \"%s\"\n", a.infoPath ? a.infoPath : "(null)");
	printf("\tinfoLen: %'d\n", a.infoLen);
	printf("}\n");
}

struct Sink {
	FILE* out = nullptr;
	size_t total = 0;
	std::vector<char> buf;
};

// writes "%d %d %d\n" per pair (cli.cu:199-206) with an integer formatter
// and one fwrite per chunk
char* put_int(char* p, int64_t v) {
	char tmp[24];
	int n = 0;
	uint64_t u = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
	do {
		tmp[n++] = (char)('0' + u % 10);
		u /= 10;
	} while (u);
	if (v < 0)
		*p++ = '-';
	while (n)
		*p++ = tmp[--n];
	return p;
}

void nn_handler(void* user, const xtn_output* o) {
	Sink* s = static_cast<Sink*>(user);
	s->total += (size_t)o->len;
	if (!s->out)
		return;
	const size_t need = (size_t)o->len * 40 + 64;
	if (s->buf.size() < need)
		s->buf.resize(need);
	char* p = s->buf.data();
	for (int64_t i = 0; i < o->len; i++) {
		p = put_int(p, o->index_pairs[i].x);
		*p++ = ' ';
		p = put_int(p, o->index_pairs[i].y);
		*p++ = ' ';
		p = put_int(p, o->pairwise_distances[i]);
		*p++ = '\n';
	}
	fwrite(s->buf.data(), 1, (size_t)(p - s->buf.data()), s->out);
}

// cli.cu:208-215: "%d %d %'lu\n" under the process locale
void overlap_handler(void* user, const xtn_output* o) {
	Sink* s = static_cast<Sink*>(user);
	s->total += (size_t)o->len;
	if (!s->out)
		return;
	for (int64_t i = 0; i < o->len; i++)
		fprintf(s->out, "%d %d %'lu\n", o->index_pairs[i].x, o->index_pairs[i].y,
		        (unsigned long)o->pairwise_frequencies[i]);
}

}  // namespace

// stage timings on stderr when XTN_CLI_TIMING is set (the bench reads them)
struct Timing {
	bool on = getenv("XTN_CLI_TIMING") != nullptr;
	std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now(), last = t0;
	std::string line;
	void mark(const char* stage) {
		if (!on)
			return;
		const auto now = std::chrono::steady_clock::now();
		char buf[64];
		snprintf(buf, sizeof(buf), " %s_ms=%.3f", stage, std::chrono::duration<double, std::milli>(now - last).count());
		line += buf;
		last = now;
	}
	~Timing() {
		if (on)
			fprintf(stderr, "xtn_cli_times%s total_ms=%.3f\n", line.c_str(),
			        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
	}
};

int main(int argc, char** argv) {
	Timing timing;
	Args args;
	setlocale(LC_ALL, "");  // cli.cu:234
	int rc = parse_args(argc, argv, &args);
	if (rc != SUCCESS)
		return rc;

	// the device is opened on a second thread while the input is parsed;
	// nothing is printed there, and a parse error still reports exactly as
	// before (the opener is joined and its handle closed)
	xtn_config cfg;
	xtn_config_init(&cfg);
	cfg.distance = args.distance;
	cfg.measure = args.measure;
	cfg.verbose = args.verbose;
	if (const char* e = getenv("XTN_MAX_RECORDS_PER_CHUNK"))
		cfg.max_records_per_chunk = atoll(e);
	xtn_handle* h = nullptr;
	int st_open = XTN_ERR_HIP;
	std::thread opener([&] { st_open = xtn_open(&h, &cfg); });
	struct Join {
		std::thread& t;
		~Join() {
			if (t.joinable())
				t.join();
		}
	} join_opener{opener};
	auto close_engine = [&] {
		if (opener.joinable())
			opener.join();
		if (h)
			xtn_close(h);
		h = nullptr;
	};

	const bool overlapMode = args.infoPath != nullptr;
	SeqStore seqs;
	seqs.maxSize = args.seqLen > 0 ? args.seqLen : 0;
	if (args.seqLen > 0) {
		seqs.chars.reserve((size_t)args.seqLen * 18);
		seqs.offsets.reserve((size_t)args.seqLen + 1);
	}
	std::vector<xtn_seq_info> info;
	if (overlapMode) {
		info.assign(args.seqLen > 0 ? (size_t)args.seqLen : 0, xtn_seq_info{0, 0, 0});
		rc = parse_info(args.infoPath, &info, args.infoLen, args.seqLen);
		if (rc != SUCCESS) {
			close_engine();
			return rc;
		}
	}
	if (args.airr) {
		rc = parse_airr_input(args.seqPath, &seqs, &info, args.seqLen, overlapMode);
	} else {
		// XTN_CLI_SERIAL_PARSE: the serial parser only (tests compare the two)
		rc = getenv("XTN_CLI_SERIAL_PARSE") ? FALLBACK
		                                     : parse_input_parallel(args.seqPath, &seqs, &info, args.seqLen, overlapMode);
		if (rc == FALLBACK) {
			seqs.chars.clear();
			seqs.offsets.assign(1, 0);
			rc = parse_input(args.seqPath, &seqs, &info, args.seqLen, overlapMode);
		}
	}
	if (rc != SUCCESS) {
		close_engine();
		return rc;
	}
	timing.mark("parse");
	if (const char* dump = getenv("XTN_CLI_PARSE_DUMP")) {
		// tests: the parsed store (offsets, residues, frequencies) and exit
		close_engine();
		FILE* f = fopen(dump, "wb");
		if (!f)
			return print_err("dump file opening failed");
		const uint32_t n = (uint32_t)seqs.size();
		fwrite(&n, sizeof(n), 1, f);
		fwrite(seqs.offsets.data(), sizeof(uint32_t), seqs.offsets.size(), f);
		fwrite(seqs.chars.data(), 1, seqs.chars.size(), f);
		for (const xtn_seq_info& x : info)
			fwrite(&x, sizeof(x), 1, f);
		fclose(f);
		return SUCCESS;
	}
	if (args.verbose)
		print_args(args);

	opener.join();
	timing.mark("open_wait");
	int st = st_open;
	if (st != XTN_OK) {
		fprintf(stderr, "Error: cannot open the GPU engine (%s)\n", xtn_status_string(st));
		return ERROR;
	}
	st = xtn_set_sequences(h, seqs.chars.data(), seqs.offsets.data(), seqs.size());
	timing.mark("upload");
	if (st != XTN_OK) {
		fprintf(stderr, "Error: %s\n", xtn_last_error(h));
		xtn_close(h);
		return ERROR;
	}
	Sink sink;
	if (args.outputPath) {
		sink.out = fopen(args.outputPath, "w");
		if (!sink.out) {
			xtn_close(h);
			return print_err("output file opening failed");
		}
	}
	st = xtn_perform(h, overlapMode ? info.data() : nullptr, overlapMode ? overlap_handler : nn_handler,
	                 &sink);
	timing.mark("perform");
	if (sink.out)
		fclose(sink.out);
	if (st != XTN_OK) {
		fprintf(stderr, "Error: %s\n", xtn_last_error(h));
		xtn_close(h);
		return ERROR;
	}
	xtn_close(h);
	timing.mark("close");
	if (args.verbose)
		printf("total output length: %'lu\n", (unsigned long)sink.total);
	return SUCCESS;
}
